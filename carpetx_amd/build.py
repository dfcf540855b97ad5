"""Build recipe for libcxp.so (hipcc, gfx950), the C++ host check program
and the CPU oracle.

    python -m carpetx_amd.build          # builds carpetx_amd/libcxp.so
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcxp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.hip", "tag.hip", "reflux.hip", "minmod_adjoint.hip", "eno_ccc.hip"]

# FMA contraction stays off so the tensor passes reproduce the reference's
# per-point arithmetic bit for bit; packed FP32 instructions are disabled.
HIPFLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    gen = [sys.executable, os.path.join(CSRC, "gen_coeffs.py"),
           os.path.join(CSRC, "cxp_coeffs.inc")]
    subprocess.check_call(gen, stdout=subprocess.DEVNULL)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)


WAVE_LIB = os.path.join(HERE, "libcxp_wave.so")


def build_wave_lib():
    """The wave harness's native RK4 stage (csrc/wave.hip), a library of its
    own beside libcxp.so."""
    tmp = WAVE_LIB + ".tmp"
    run([HIPCC] + HIPFLAGS + ["-shared", "-o", tmp, os.path.join(CSRC, "wave.hip")])
    os.replace(tmp, WAVE_LIB)


HOST_CHECK = os.path.join(HERE, "host", "prolongate_check")


def build_host_check():
    """The C++ host API's check program, linked against libcxp.so."""
    tmp = HOST_CHECK + ".tmp"
    run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
         "-Wno-unused-command-line-argument",
         "-I" + os.path.join(ROOT, "include"),
         os.path.join(HERE, "host", "prolongate_check.cpp"),
         "-L" + HERE, "-lcxp", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, HOST_CHECK)


INTERP_CHECK = os.path.join(ROOT, "tests", "amrex_shim", "interpolater_check")


def build_interpolater_check():
    """The Interpolater adapter (host/cxp_interpolater.hxx) compiled against
    the AMReX / Cactus shim of tests/amrex_shim, linked to libcxp.so."""
    tmp = INTERP_CHECK + ".tmp"
    shim = os.path.dirname(INTERP_CHECK)
    run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
         "-Wno-unused-command-line-argument",
         "-I" + shim, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host"),
         os.path.join(shim, "interpolater_check.cpp"),
         "-L" + HERE, "-lcxp", "-Wl,-rpath,$ORIGIN/../../carpetx_amd", "-o", tmp])
    os.replace(tmp, INTERP_CHECK)


MINMOD_HOST = os.path.join(HERE, "host", "libminmod_adjoint_host.so")


def build_minmod_host():
    """The MINMOD adjoint's record and gather routines built for the host
    with bounds-checked accessors (host/minmod_adjoint_host.cpp): the CPU
    check of the kernels' indexing, loaded by the tests only."""
    tmp = MINMOD_HOST + ".tmp"
    run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
         "-ffp-contract=off", "-fPIC", "-shared",
         "-Wno-unused-command-line-argument",
         os.path.join(HERE, "host", "minmod_adjoint_host.cpp"), "-o", tmp])
    os.replace(tmp, MINMOD_HOST)


ADJ_ZF_HOST = os.path.join(HERE, "host", "libadj_zfirst_host.so")


def build_adj_zfirst_host():
    """The z-first linear adjoint's routines built for the host with checked
    accesses (host/adj_zfirst_host.cpp): the CPU check of the kernel's
    indexing, loaded by the tests only."""
    tmp = ADJ_ZF_HOST + ".tmp"
    run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
         "-ffp-contract=off", "-fPIC", "-shared",
         "-Wno-unused-command-line-argument",
         os.path.join(HERE, "host", "adj_zfirst_host.cpp"), "-o", tmp])
    os.replace(tmp, ADJ_ZF_HOST)


ENO_HOST = os.path.join(HERE, "host", "libeno_ccc_host.so")


def build_eno_host():
    """The CCC ENO routines built for the host with checked accesses
    (host/eno_ccc_host.cpp): the CPU check of the kernels' indexing, loaded
    by the tests only."""
    tmp = ENO_HOST + ".tmp"
    run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
         "-ffp-contract=off", "-fPIC", "-shared",
         "-Wno-unused-command-line-argument",
         os.path.join(HERE, "host", "eno_ccc_host.cpp"), "-o", tmp])
    os.replace(tmp, ENO_HOST)


def build_oracle():
    run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])


def main():
    t0 = time.time()
    build_lib()
    build_wave_lib()
    build_host_check()
    build_minmod_host()
    build_eno_host()
    build_adj_zfirst_host()
    build_interpolater_check()
    print("built %s in %.1fs" % (LIB, time.time() - t0))


if __name__ == "__main__":
    main()
