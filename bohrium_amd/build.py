"""Build the native libraries of bohrium_amd in-tree (bohrium_amd/lib/).

    libbh.so            core IR, config, component loader, host memory, jitk fusion (g++)
    libbhc.so           the bhc C-ABI bridge runtime (g++)
    libbh_vem_node.so   the pass-through node manager component (g++)
    libbh_ve_hip.so     the hip vector engine component + bhhip C-ABI (hipcc, gfx950)
    libbh_blas_hip.so   blas_* extension methods on the matrix cores (hipcc, gfx950)

Every call recompiles; objects go to bohrium_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(ROOT, "lib")
OBJ = os.path.join(ROOT, "_build")
INC = os.path.join(CSRC, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + INC]
# HIP sources: gfx950 only, packed-fp32 instructions disabled in device code.
HIPFLAGS = ["-std=c++17", "-O2", "-fPIC", "-I" + INC, "--offload-arch=gfx950",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]

CORE = ["core/bh_type.cpp", "core/bh_constant.cpp", "core/bh_view.cpp", "core/bh_instruction.cpp", "core/bh_ir.cpp",
        "core/bh_opcode.cpp", "core/bh_config_parser.cpp", "core/bh_component.cpp",
        "core/bh_main_memory.cpp", "core/bh_mem_signal.cpp", "jitk/block.cpp", "jitk/fuser.cpp"]
BRIDGE = ["bridge/runtime.cpp"]
NODE = ["vem_node/node.cpp"]
VE_HIP_CXX = ["ve_hip/codegen.cpp", "ve_hip/rewrite.cpp", "ve_hip/features.cpp", "ve_hip/tuner.cpp"]
VE_HIP_HIP = ["ve_hip/bhhip.cpp", "ve_hip/engine.cpp", "ve_hip/batch.cpp", "ve_hip/costmodel.hip"]
BLAS_HIP = ["ext_blas/blas_hip.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "_").rsplit(".", 1)[0] + ".o")


def _compile(src, hip=False):
    out = _obj(src)
    flags = HIPFLAGS + (["-O3"] if src in BLAS_HIP else [])
    cmd = ([HIPCC] + flags if hip else ["g++"] + CXXFLAGS) + ["-c", os.path.join(CSRC, src), "-o", out]
    _run(cmd)
    return out


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    jobs = [(s, False) for s in CORE + BRIDGE + NODE + VE_HIP_CXX] + [(s, True) for s in VE_HIP_HIP + BLAS_HIP]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = dict(zip([s for s, _ in jobs], ex.map(lambda j: _compile(*j), jobs)))
    rpath = "-Wl,-rpath,$ORIGIN"
    libbh = os.path.join(LIB, "libbh.so")
    _run(["g++", "-shared", "-o", libbh] + [objs[s] for s in CORE] + ["-ldl"])
    _run(["g++", "-shared", "-o", os.path.join(LIB, "libbhc.so")] + [objs[s] for s in BRIDGE]
         + ["-L" + LIB, "-lbh", rpath, "-ldl"])
    _run(["g++", "-shared", "-o", os.path.join(LIB, "libbh_vem_node.so")] + [objs[s] for s in NODE]
         + ["-L" + LIB, "-lbh", rpath])
    _run([HIPCC, "-shared", "-fPIC", "-o", os.path.join(LIB, "libbh_ve_hip.so")]
         + [objs[s] for s in VE_HIP_CXX + VE_HIP_HIP]
         + ["-L" + LIB, "-lbh", rpath, "-L" + os.path.join(ROCM, "lib"), "-lhiprtc", "-lamdhip64",
            "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    _run([HIPCC, "-shared", "-fPIC", "-o", os.path.join(LIB, "libbh_blas_hip.so")] + [objs[s] for s in BLAS_HIP]
         + ["-L" + LIB, "-lbh", rpath, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
            "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    # CPython extension with the per-operation bridge calls.
    import sysconfig
    pyinc = sysconfig.get_paths()["include"]
    _run(["g++"] + CXXFLAGS + ["-shared", "-I" + pyinc, os.path.join(CSRC, "bridge", "pybridge.cpp"),
                               "-o", os.path.join(ROOT, "_bhc.so"), "-L" + LIB, "-lbhc", "-Wl,-rpath,$ORIGIN/lib"])
    if verbose:
        print("built:", sorted(os.listdir(LIB)) + ["_bhc.so"])


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
