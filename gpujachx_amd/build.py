"""Build the native parts in-tree.

    python -m gpujachx_amd.build [--no-ref]

Produces
  gpujachx_amd/lib/libhypjac.so   HIP library for gfx950 (kernels + C ABI)
  gpujachx_amd/lib/strat.so       strategy plugin (reference symbol ABI)
  gpujachx_amd/lib/x1             command-line driver (HDF5), if libhdf5 is found
  gpujachx_amd/lib/h5gen, h5check HDF5 input writer and accuracy verifier
  oracle/libhjoracle.so           CPU oracle (test infrastructure)
  oracle/_ref/strat.so            the reference's own tables, if /root/reference exists
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpujachx_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: no packed FP32 instructions; -ffp-contract=off: only the
# fma the code writes out.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HDF5_PREFIXES = ["/opt/conda", "/usr", "/usr/local"]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def _hdf5_prefix():
    for p in HDF5_PREFIXES:
        if (os.path.exists(os.path.join(p, "include", "hdf5.h"))
                and os.path.exists(os.path.join(p, "lib", "libhdf5_hl.so"))):
            return p
    return None


def build_library():
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in ["hj_kernels.hip", "hj_qr.hip", "hj_mlp.hip", "hj_batch.hip", "hj_wide.hip", "hj_host.cpp",
                "hj_strat.cpp"]:
        obj = os.path.join(LIB, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libhypjac.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_strat():
    gen = os.path.join(LIB, "hj_stratgen")
    _run(["g++", "-O2", "-std=c++17", "-o", gen, os.path.join(CSRC, "hj_stratgen.cpp"),
          os.path.join(CSRC, "hj_strat.cpp"), "-ldl"])
    csrc = os.path.join(LIB, "strat_tables.c")
    _run([gen, csrc])
    out = os.path.join(LIB, "strat.so")
    _run(["gcc", "-O0", "-fPIC", "-shared", "-o", out + ".tmp", csrc])
    os.replace(out + ".tmp", out)
    os.remove(csrc)
    return out


def build_tools():
    pre = _hdf5_prefix()
    if pre is None:
        print("libhdf5 not found: x1/h5gen/h5check not built", flush=True)
        return
    inc, lib = os.path.join(pre, "include"), os.path.join(pre, "lib")
    common = ["-O2", "-std=c++17", "-I" + inc, "-I" + os.path.join(ROOT, "include")]
    # The HDF5 prefix may also hold an older libstdc++ than the HIP runtime
    # needs, so the tools do not search it: a private directory of links to
    # the two HDF5 libraries is put on their run path instead.
    h5dir = os.path.join(LIB, "hdf5")
    os.makedirs(h5dir, exist_ok=True)
    for stem in ("libhdf5", "libhdf5_hl"):
        real = os.path.realpath(os.path.join(lib, stem + ".so"))
        soname = subprocess.check_output(["objdump", "-p", real], text=True)
        soname = [l.split()[-1] for l in soname.splitlines() if "SONAME" in l][0]
        for name in (stem + ".so", soname):
            dst = os.path.join(h5dir, name)
            if os.path.lexists(dst):
                os.remove(dst)
            os.symlink(real, dst)
    link_h5 = ["-L" + h5dir, "-Wl,-rpath,$ORIGIN/hdf5", "-lhdf5_hl", "-lhdf5"]
    srcs = {
        "x1": ["x1_main.cpp"],
        "h5gen": ["h5gen_main.cpp"],
        "h5check": ["h5check_main.cpp"],
    }
    for name, files in srcs.items():
        paths = [os.path.join(CSRC, f) for f in files]
        if not all(os.path.exists(p) for p in paths):
            continue
        extra = []
        if name == "x1":
            extra = ["-L" + LIB, "-Wl,-rpath,$ORIGIN", "-lhypjac"]
        if name == "h5check":
            extra = ["-fopenmp"]
        _run(["g++"] + common + ["-o", os.path.join(LIB, name)] + paths + extra + link_h5)


def build_pmc():
    """hj_pmc: the profiling driver without Python (counter runs)."""
    out = os.path.join(LIB, "hj_pmc")
    _run([HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-o", out,
          os.path.join(CSRC, "pmc_main.cpp"), "-L" + LIB, "-Wl,-rpath,$ORIGIN", "-lhypjac"])
    return out


def build_oracle(ref=True):
    out = os.path.join(ORACLE, "libhjoracle.so")
    _run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-math-errno", "-fopenmp", "-fPIC",
          "-shared", "-Wall", "-o", out + ".tmp", os.path.join(ORACLE, "hjoracle.c"),
          os.path.join(ORACLE, "hjoracle_qr.c"), os.path.join(ORACLE, "hjoracle_wide.c"), "-lm"])
    os.replace(out + ".tmp", out)
    if ref and os.path.isdir("/root/reference/strat"):
        _run(["bash", os.path.join(ORACLE, "build_ref.sh"), "/root/reference"])
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_oracle(ref="--no-ref" not in argv)
    build_library()
    build_pmc()
    build_strat()
    build_tools()


if __name__ == "__main__":
    main()
