"""Build libawig.so in-tree with hipcc for gfx950 (``python -m accuratewigner_amd.build``)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(HERE, "csrc", f) for f in
           ("awig_fft.hip", "awig_pass_f32.hip", "awig_pass_f64.hip", "awig_pass_c64.hip", "awig_scratch.hip",
                      "awig_api.cpp")]
TARGET = os.path.join(HERE, "libawig.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 ops disabled for all HIP code on this target
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _clean(stderr):
    return "\n".join(l for l in stderr.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=False):
    """Compile each translation unit in parallel, then link libawig.so (atomically replaced)."""
    import tempfile
    tmpdir = tempfile.mkdtemp(prefix="awig_build_")
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for src, p in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("hipcc failed on %s (%d):\n%s" % (src, p.returncode, _clean(err)))
    tmp = TARGET + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-lhiprtc", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed (%d):\n%s" % (res.returncode, _clean(res.stderr)))
    os.replace(tmp, TARGET)
    for o in objs:
        os.remove(o)
    os.rmdir(tmpdir)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
