"""Build libtkfmi.so in-tree (hipcc --offload-arch=gfx950)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(_HERE), "csrc")


def build(arch="gfx950", force=True, verbose=False):
    """Compile the HIP sources into tkfmi/_lib/libtkfmi.so.  Always
    recompiles when ``force`` (the build is the 'does it build' check)."""
    env = dict(os.environ)
    env.setdefault("HIPCC", "/opt/rocm/bin/hipcc")
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, env=env,
                       stdout=subprocess.DEVNULL)
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    r = subprocess.run(["make", "-C", CSRC, f"-j{jobs}", f"ARCH={arch}"], env=env,
                       capture_output=True, text=True)
    if verbose or r.returncode != 0:
        print(r.stdout[-4000:])
        print(r.stderr[-4000:])
    if r.returncode != 0:
        raise RuntimeError("libtkfmi build failed")
    return os.path.join(_HERE, "_lib", "libtkfmi.so")
