"""Build the network libraries: one libkrome_<network>.so per configured
network, hand-written HIP for gfx950 compiled with hipcc (in-tree).

    python3 krome-x_amd/build.py [--regen]

--regen regenerates the network headers from the .knet files first (needs
the KROME-X network folder, KROME_NETWORKS or /root/reference/networks).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
GEN = CSRC / "generated"
NETWORKS = ("primordial", "primordial3", "galaxy_ism", "primordialZ",
            "galaxy_ism_heat", "cloud", "NO", "SM", "coolHCN", "lowmetal",
            "primordial2", "primordial_wD", "primordial_photo",
            "primordial_photoH2", "auto", "chianti", "primordial_UV",
            "xrays", "primordialZ_UV", "photobin")
# generated networks built only on request (build(net)): ions and ionsH2
# (55 and 58 species, 45 and 48 photo reactions) take about ten minutes of
# hipcc each for an 8-10 MB library; their generated code is checked through
# the oracle, which runs the same headers
ON_REQUEST = ("ions", "ionsH2")
# test builds of a network with other compile-time forms: name -> (network,
# extra flags); primordial_rhsloop evaluates the rhs and the species
# Jacobian by the loop form over index tables (the form of networks above
# netgen.RHS_LOOP_MIN reactions) instead of the unrolled code
VARIANTS = {"primordial_rhsloop": ("primordial", ["-DKX_FORCE_RHS_LOOP"])}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # no packed FP32 instructions (v_pk_*_f32) in any kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]
# a library's optimisation level where it differs from FLAGS' -O3 (the
# later flag wins), each measured with results bitwise equal:
# primordial3's lockstep mixed kernel on the 2^20 collapse cells of the
# bench ran 934.9 ms at -O2 against 965.9 ms at -O3; primordial's lockstep
# kernel 291 against 297-308 ms on 2^18 fixed-Tgas cells and 110.8 against
# 113.5-114.4 ms on 4096 (its plain kernel unchanged); the other libraries
# are not measured at -O2
OPT = {"primordial3": "-O2", "primordial": "-O2"}


def lib_path(net):
    return HERE / ("libkrome_%s.so" % net)


# libraries of several networks in one persistent launch (csrc/kx_multi.h
# over a file generated here): name -> (the networks in the C interface's
# order, each with Tgas evolved or fixed; the launch order, the costliest
# cells first)
MULTI = {
    "multi": ((("primordial", False), ("primordial3", True)),
              ("primordial3", "primordial")),
    # the uneven batch: the lockstep machine for primordial, the lane queue
    # for primordialZ (27 equations) and cloud (452)
    "multi_c3": ((("primordial", False), ("primordialZ", False),
                  ("cloud", False)),
                 ("cloud", "primordialZ", "primordial")),
}


def multi_source(name):
    """write csrc/generated/<name>.hip, the network list of a multi
    library, and return its path"""
    nets, order = MULTI[name]
    index = {n: i for i, (n, _) in enumerate(nets)}
    et = dict(nets)
    lines = ["// generated by krome-x_amd/build.py: the networks of the multi",
             "// library %s (C-interface index, name, Tgas evolved), in launch"
             % name, "// order"]
    lines += ['#include "net_%s.h"' % n for n in order]
    lines += ['#define KX_MULTI_NAME "%s"' % name,
              "#define KX_MULTI_COUNT %d" % len(nets),
              "#define KX_MULTI_NETS(X) %s" % " ".join(
                  "X(%d, %s, %s)" % (index[n], n, "true" if et[n] else "false")
                  for n in order),
              '#include "kx_multi.h"', ""]
    path = GEN / ("%s.hip" % name)
    text = "\n".join(lines)
    if not path.exists() or path.read_text() != text:
        path.write_text(text)
    return path


def regen(networks_dir=None):
    nd = networks_dir or os.environ.get("KROME_NETWORKS",
                                        "/root/reference/networks")
    if not Path(nd).is_dir():
        return False
    sys.path.insert(0, str(HERE))
    import netgen
    netgen.main(["--networks", str(nd), "--out", str(GEN)])
    return True


def _command(net, variant, tmp):
    if net in MULTI:
        return [HIPCC, *FLAGS, "-I", str(CSRC), "-I", str(GEN),
                str(multi_source(net)), "-o", str(tmp)]
    extra = VARIANTS[variant][1] if variant else []
    opt = [OPT[variant or net]] if (variant or net) in OPT else []
    return [HIPCC, *FLAGS, *opt, *extra, "-I", str(CSRC), "-I", str(GEN),
            "-DKX_NET=net_%s" % net,
            '-DKX_NET_HEADER="net_%s.h"' % net,
            '-DKX_META_HEADER="net_%s_meta.h"' % net,
            str(CSRC / "kx_abi.hip"), "-o", str(tmp)]


def build(net, verbose=False, variant=None):
    out = lib_path(variant or net)
    tmp = out.with_suffix(".so.tmp")
    cmd = _command(net, variant, tmp)
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_all(verbose=False, regen_headers=False):
    """every network library and test variant, compiled concurrently"""
    if regen_headers:
        regen()
    jobs = ([(n, None) for n in NETWORKS] + [(net, v) for v, (net, _)
                                             in VARIANTS.items()]
            + [(m, None) for m in MULTI])
    procs = []
    for net, variant in jobs:
        out = lib_path(variant or net)
        tmp = out.with_suffix(".so.tmp")
        cmd = _command(net, variant, tmp)
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd), tmp, out))
    failed = [p.args for p, _, _ in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    for _, tmp, out in procs:
        os.replace(tmp, out)
    return [out for _, _, out in procs]

if __name__ == "__main__":
    build_all(verbose=True, regen_headers="--regen" in sys.argv)
