"""The bindings INTEGRATION.md shows a maintainer are compiled and linked
against libewfa, so that the documented code cannot drift from ewfa.h."""
import os
import re
import shutil
import subprocess

import pytest

import ewfa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")
INCLUDE = os.path.join(ROOT, "include")


def blocks(lang):
    text = open(DOC).read()
    return re.findall(r"```" + lang + r"\n(.*?)```", text, flags=re.S)


@pytest.mark.parametrize("lang,compiler,std", [("cpp", "g++", "-std=c++17"), ("c", "gcc", "-std=c11")])
def test_doc_bindings_compile_and_link(tmp_path, lang, compiler, std):
    if not shutil.which(compiler):
        pytest.skip(f"{compiler} not found")
    found = blocks(lang)
    assert found, f"no ```{lang} block in INTEGRATION.md"
    for i, code in enumerate(found):
        src = os.path.join(tmp_path, f"snippet{i}.{lang}")
        with open(src, "w") as f:
            f.write(code)
        out = os.path.join(tmp_path, f"libsnippet{i}.so")
        # -z defs: every symbol the snippet uses must resolve in libewfa
        cmd = [compiler, std, "-Wall", "-Werror", "-Wno-unused-function", "-fPIC", "-shared", "-I", INCLUDE,
               src, "-o", out, "-L", ewfa.BUILD_DIR, "-lewfa", "-Wl,-z,defs"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
