// args.cpp -- the wfe.aligner command-line contract.
//
// The option table is the reference's (src/aligner.cu:75-139), parsed with
// the reference's rules (utils/arg_handler.c:31-138): argv is scanned from
// index 0; "--name" matches a long option exactly, "-x..." matches the first
// option whose short letter is x; an option that takes a value consumes the
// next argument whatever it looks like (integers through atoll); unknown
// arguments are ignored; parsing fails only when a required option is
// missing.  Five options are additions and come after the reference's six.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ewfa.h"

namespace {

enum ArgType { ARG_INT, ARG_STR, ARG_NO_VALUE };

struct Option {
    const char* name;
    const char* description;
    char short_arg;
    const char* long_arg;
    bool required;
    ArgType type;
};

// The first six rows are the reference's table, text for text.
const Option kOptions[] = {
    {"Sequences file", "File containing the sequences to align.", 'f', "file", true, ARG_STR},
    {"Number of alignments",
     "Number of alignments to read from the file (default=all alignments)", 'n',
     "num-alignments", true, ARG_INT},
    {"Sequence length", "Maximum sequence length.", 'l', "seq-len", true, ARG_INT},
    {"Batch size", "Number of alignments per batch (default=num-alignments).", 'b', "batch-size",
     false, ARG_INT},
    {"Number of CPU threads",
     "Number of CPU threads, each CPU thread creates two streams to overlap compute and memory "
     "transfers. (default=1)",
     't', "cpu-threads", false, ARG_INT},
    {"Print CIGARS", "Print CIGARS to stdout", 'p', "print-cigars", false, ARG_NO_VALUE},
    // additions
    {"Max distance",
     "Stop the search at this distance like the reference's cap (it uses 64); default: exact "
     "for every distance.",
     'm', "max-distance", false, ARG_INT},
    {"Check CIGARS", "Validate every CIGAR on the host and report Correct alignments: a/b.", 'c',
     "check", false, ARG_NO_VALUE},
    {"Standard CIGAR", "Print standard run-length CIGARs (=, X, I, D) instead of the reference text.",
     's', "standard-cigar", false, ARG_NO_VALUE},
    {"Plain launch", "Use the reference's per-batch launch model instead of the persistent kernel.",
     'P', "plain", false, ARG_NO_VALUE},
    {"Launch policy", "Policy file written by the autotuner for this workload (default: the built-in table).",
     'u', "policy", false, ARG_STR},
};
constexpr int kNumOptions = (int)(sizeof kOptions / sizeof kOptions[0]);

struct Parsed {
    bool parsed;
    int64_t int_val;
    const char* str_val;
};

}  // namespace

extern "C" {

int ewfa_cli_parse(int argc, char** argv, ewfa_cli_args_t* out) {
    Parsed st[kNumOptions] = {};
    int cur = -1;
    bool read_value = false;
    for (int i = 0; i < argc; ++i) {
        const char* a = argv[i];
        if (!read_value) {
            const size_t len = strlen(a);
            if (a[0] != '-') continue;
            cur = -1;
            if (len >= 2 && a[1] == '-') {
                for (int o = 0; o < kNumOptions; ++o)
                    if (!strcmp(kOptions[o].long_arg, a + 2)) {
                        cur = o;
                        break;
                    }
            } else {
                for (int o = 0; o < kNumOptions; ++o)
                    if (kOptions[o].short_arg == a[1]) {
                        cur = o;
                        break;
                    }
            }
            if (cur < 0) continue;
            if (kOptions[cur].type != ARG_NO_VALUE) read_value = true;
            else st[cur].parsed = true;
        } else {
            if (kOptions[cur].type == ARG_INT) st[cur].int_val = atoll(a);
            else st[cur].str_val = a;
            st[cur].parsed = true;
            read_value = false;
            cur = -1;
        }
    }
    if (out) {
        memset(out, 0, sizeof *out);
        out->file = st[0].str_val;
        out->num_alignments = st[1].int_val;
        out->seq_len = st[2].int_val;
        out->batch_size = st[3].int_val;
        out->batch_set = st[3].parsed;
        out->cpu_threads = st[4].int_val;
        out->threads_set = st[4].parsed;
        out->print_cigars = st[5].parsed;
        out->max_distance = st[6].parsed ? st[6].int_val : 0;
        out->check = st[7].parsed;
        out->standard_cigar = st[8].parsed;
        out->plain = st[9].parsed;
        out->policy = st[10].parsed ? st[10].str_val : nullptr;
    }
    for (int o = 0; o < kNumOptions; ++o)
        if (kOptions[o].required && !st[o].parsed) return 0;
    return 1;
}

// utils/arg_handler.c:140-187
void ewfa_cli_usage(void) {
    fprintf(stderr, "Options:\n");
    for (int o = 0; o < kNumOptions; ++o) {
        const Option& op = kOptions[o];
        const char* type_str = op.type == ARG_INT ? "int" : "string";
        if (op.type == ARG_NO_VALUE) {
            if (op.required)
                fprintf(stderr, "\t-%c, --%-30s(required) %s: %s\n", op.short_arg, op.long_arg, op.name,
                        op.description);
            else
                fprintf(stderr, "\t-%c, --%-30s%s: %s\n", op.short_arg, op.long_arg, op.name,
                        op.description);
        } else {
            if (op.required)
                fprintf(stderr, "\t-%c, --%-30s(%s, required) %s: %s\n", op.short_arg, op.long_arg,
                        type_str, op.name, op.description);
            else
                fprintf(stderr, "\t-%c, --%-30s(%s) %s: %s\n", op.short_arg, op.long_arg, type_str,
                        op.name, op.description);
        }
    }
}

}  // extern "C"
