"""Builds librawspec.so (C host layer + gfx950 HIP kernels) in-tree.

    python -m rawspec_amd.build          # compile and link
    python -m rawspec_amd.build --check  # also print the exported symbols

The library lands at rawspec_amd/librawspec.so.  Objects go to
rawspec_amd/_build/.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librawspec.so")
CLI_SRC = os.path.join(CSRC, "cli")
CLI = os.path.join(HERE, "bin", "rawspec")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

ARCH = os.environ.get("RAWSPEC_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_SOURCES = ["rsd_runtime.hip", "k_small.hip", "k_fourstep.hip", "k_mfma.hip", "k_misc.hip"]
C_SOURCES = ["rawspec_api.c", "raw_header.c", "fb_header.c"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INC}", f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           f"-I{INC}", f"-I{CSRC}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


SAN = os.path.join(HERE, "_san")


def san_flags(sanitize):
    """gcc flags of a sanitizer build ("address,undefined" or "thread"):
    every finding aborts the process."""
    if not sanitize:
        return []
    return [f"-fsanitize={sanitize}", "-fno-omit-frame-pointer", "-g",
            "-fno-sanitize-recover=all"]


def san_libs(sanitize):
    """gcc's sanitizer runtimes, linked into a library whose host C is
    instrumented (the HIP objects are not)."""
    libs = []
    for kind, name in (("address", "libasan.so"), ("undefined", "libubsan.so"),
                       ("thread", "libtsan.so")):
        if sanitize and kind in sanitize.split(","):
            libs.append(subprocess.run([CC, f"-print-file-name={name}"], capture_output=True,
                                       text=True, check=True).stdout.strip())
    return libs


RESOURCE_FIELDS = ("VGPRs", "AGPRs", "ScratchSize [bytes/lane]", "Occupancy [waves/SIMD]")


def kernel_resources(remarks: str):
    """(kernel, {field: int}) per kernel from the compiler's
    kernel-resource-usage remarks."""
    out, cur = [], None
    for line in remarks.splitlines():
        if "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split("[-Rpass")[0].strip()
        if body.startswith("Function Name:"):
            cur = (body.split(":", 1)[1].strip(), {})
            out.append(cur)
        elif cur is not None and ":" in body:
            key, val = body.rsplit(":", 1)
            if key.strip() in RESOURCE_FIELDS:
                cur[1][key.strip()] = int(val.strip())
    return out


def compile_one(src, build_dir=BUILD, defines=(), sanitize=None):
    """Compiles one source into build_dir; returns (object, per-kernel
    resources of a HIP source)."""
    path = os.path.join(CSRC, src)
    out = os.path.join(build_dir, os.path.basename(src) + ".o")
    dflags = [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", *dflags, "-x",
               "hip", path, "-o", out]
    else:
        cmd = [CC, "-c", *C_FLAGS, *san_flags(sanitize), *dflags, path, "-o", out]
    r = _run(cmd)
    return out, (kernel_resources(r.stderr) if src.endswith(".hip") else [])


def build(jobs: int | None = None, verbose: bool = False, out: str = LIB,
          defines=(), build_dir: str = BUILD, sanitize: str | None = None) -> str:
    """Compile every source and link librawspec.so; returns its path.

    defines: extra preprocessor definitions (the kernels' compile-time
    tunables, e.g. "RSK_E_BIG=8"); out/build_dir redirect the result;
    sanitize: instrument the host C (gcc -fsanitize=...; see sanitize.sh
    beside the tests)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(build_dir, exist_ok=True)
    srcs = HIP_SOURCES + C_SOURCES
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        done = list(ex.map(lambda s: compile_one(s, build_dir, defines, sanitize), srcs))
    objs = [o for o, _ in done]
    # registers, scratch and occupancy of every kernel, as the compiler
    # reports them (the tests check that no kernel needs scratch memory,
    # which the runtime would allocate outside rawspec_amd_plan's figure)
    with open(os.path.join(build_dir, "kernel_resources.tsv"), "w") as f:
        f.write("kernel\t" + "\t".join(RESOURCE_FIELDS) + "\n")
        for _, res in done:
            for name, fields in res:
                f.write(name + "\t" + "\t".join(str(fields.get(k, -1)) for k in RESOURCE_FIELDS)
                        + "\n")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs,
          *san_libs(sanitize), "-Wl,--no-undefined", "-Wl,-soname,librawspec.so"])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


CLI_SOURCES = ["cli_main.c", "cli_ingest.c", "cli_output.c", "cli_net.c"]
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def build_fbh5_plugin(out_dir: str, verbose: bool = False):
    """librawspec_fbh5.so, the FBH5 writer the program loads for -j; built
    only where the HDF5 headers and libraries (with the high-level library)
    are found.  Returns its path or None."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and
            os.path.exists(os.path.join(libdir, "libhdf5_hl.so"))):
        if verbose:
            print(f"no HDF5 under {HDF5_DIR}: the program is built without FBH5 output")
        return None
    obj = os.path.join(BUILD, "cli_fbh5_writer.o")
    out = os.path.join(out_dir, "librawspec_fbh5.so")
    _run([CC, "-c", *C_FLAGS, f"-I{CLI_SRC}", f"-I{inc}", os.path.join(CLI_SRC, "fbh5_writer.c"),
          "-o", obj])
    _run([CC, "-shared", "-o", out + ".tmp", obj, f"-L{libdir}", "-lhdf5_hl", "-lhdf5",
          f"-Wl,-rpath,{libdir}"])
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


def build_cli(lib: str = LIB, out: str = CLI, verbose: bool = False, sanitize: str | None = None,
              build_dir: str = BUILD) -> str:
    """Compile the rawspec program (host C) against the built library; it
    finds librawspec.so next to its own directory at run time."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    for src in CLI_SOURCES:
        o = os.path.join(build_dir, "cli_" + src + ".o")
        _run([CC, "-c", *C_FLAGS, *san_flags(sanitize), "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", f"-I{CLI_SRC}", os.path.join(CLI_SRC, src), "-o", o])
        objs.append(o)
    tmp = out + ".tmp"
    _run([CC, *san_flags(sanitize), "-o", tmp, *objs, f"-L{os.path.dirname(lib)}", "-lrawspec",
          "-Wl,-rpath,$ORIGIN/..", f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib",
          "-lpthread", "-ldl", "-lm"])
    os.replace(tmp, out)
    if not sanitize:
        build_fbh5_plugin(os.path.dirname(out), verbose)
    if verbose:
        print(f"built {out}")
    return out


def exported_symbols(path: str = LIB):
    nm = shutil.which("nm") or "nm"
    r = _run([nm, "-D", "--defined-only", path])
    return sorted(line.split()[-1] for line in r.stdout.splitlines() if " T " in line)


def build_sanitized(sanitize: str, verbose: bool = False):
    """librawspec.so and the rawspec program with gcc's sanitizers on the
    host C, under rawspec_amd/_san/<first kind>/ (the normal build is left
    alone); returns (library, program)."""
    d = os.path.join(SAN, sanitize.split(",")[0])
    lib = build(verbose=verbose, out=os.path.join(d, "librawspec.so"),
                build_dir=os.path.join(d, "_build"), sanitize=sanitize)
    cli = build_cli(lib=lib, out=os.path.join(d, "bin", "rawspec"), verbose=verbose,
                    sanitize=sanitize, build_dir=os.path.join(d, "_build"))
    return lib, cli


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--sanitize", default=None,
                    help="address,undefined or thread: an instrumented build under rawspec_amd/_san/")
    a = ap.parse_args(argv)
    if a.sanitize:
        build_sanitized(a.sanitize, verbose=True)
        return 0
    build(a.j, verbose=True)
    build_cli(verbose=True)
    if a.check:
        for s in exported_symbols():
            if s.startswith(("rawspec", "get_")):
                print(s)


if __name__ == "__main__":
    sys.exit(main())
