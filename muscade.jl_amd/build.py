"""Build libmuscade_hip.so for gfx950 (hipcc) in-tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", "mh_lib.hip")]
OUT = os.path.join(HERE, "muscade_amd", "libmuscade_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


# rocSOLVER and rocBLAS: the dense block LU of DirectXUA's all-steps system (mh_big_solve)
LIBS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]


def _sources_digest():
    """Digest of every source the library is built from: hipcc reads them once for the device
    pass and again for the host pass, so a change in between gives host launches of kernels
    that the code object does not hold."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(HERE, "csrc")
    inc = os.path.join(HERE, "..", "include")
    for d in (csrc, inc):
        for f in sorted(os.listdir(d)):
            with open(os.path.join(d, f), "rb") as fh:
                h.update(f.encode() + fh.read())
    return h.hexdigest()


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + SRC + LIBS
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    before = _sources_digest()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-6000:])
    if _sources_digest() != before:
        os.remove(OUT + ".tmp")
        raise RuntimeError("the sources changed while hipcc ran: build again")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
