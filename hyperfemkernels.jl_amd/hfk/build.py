"""Build libhfk.so for gfx950 in-tree (next to this file) with hipcc.  No cmake: one hipcc line per source,
compiled in parallel, then one link line.  ``python -m hfk.build`` (from ``hyperfemkernels.jl_amd/``) or
``build()`` from Python."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
LIB = os.path.join(HERE, "libhfk.so")
SOURCES = ["hfk_kernels.hip", "hfk_capi.hip", "hfk_solver.hip", "hfk_pmg.hip", "hfk_net.hip", "hfk_oseen.hip",
           "hfk_rtc.hip", "hfk_skel.hip", "hfk_sys.hip", "hfk_pattern.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions stay off (-packed-fp32-ops): the compiler may not form v_pk_* FP32 arithmetic.
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{INCLUDE}"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


# device headers compiled again at run time by hiprtc (laws given as source): embedded in the library as text
EMBEDDED = (("kTypesSrc", "hfk_types.hpp"), ("kPointSrc", "hfk_point.hpp"), ("kOseenFormSrc", "hfk_oseen_form.hpp"),
            ("kSkelSrc", "hfk_skel.hpp"))


def write_embedded(objdir: str) -> None:
    parts = ["// generated by hfk/build.py from the device headers; do not edit\n"]
    for var, name in EMBEDDED:
        text = open(os.path.join(CSRC, name)).read()
        if ")HFKSRC\"" in text:
            raise RuntimeError(f"{name} contains the raw-string delimiter")
        parts.append(f'static const char {var}[] = R"HFKSRC({text})HFKSRC";\n')
    path = os.path.join(objdir, "hfk_embedded.inc")
    new = "".join(parts)
    if not os.path.exists(path) or open(path).read() != new:
        with open(path, "w") as f:
            f.write(new)


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    write_embedded(objdir)
    objs, cmds = [], []
    for s in SOURCES:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC, *FLAGS, f"-I{objdir}", "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-L/opt/rocm/lib", "-lhiprtc", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
