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
            "galaxy_ism_heat", "NO", "SM", "coolHCN", "lowmetal",
            "primordial2", "primordial_wD", "primordial_photo",
            "primordial_photoH2", "auto", "chianti", "primordial_UV",
            "xrays", "primordialZ_UV", "photobin")
# generated networks built only on request (build(net)): ions and ionsH2
# (55 and 58 species, 45 and 48 photo reactions) take about ten minutes of
# hipcc each for a 7 MB library; their generated code is checked through
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


def lib_path(net):
    return HERE / ("libkrome_%s.so" % net)


# the library of several networks in one persistent launch (kx_multi.hip)
MULTI = "multi"


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
    if net == MULTI:
        return [HIPCC, *FLAGS, "-I", str(CSRC), "-I", str(GEN),
                str(CSRC / "kx_multi.hip"), "-o", str(tmp)]
    extra = VARIANTS[variant][1] if variant else []
    return [HIPCC, *FLAGS, *extra, "-I", str(CSRC), "-I", str(GEN),
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
            + [(MULTI, None)])
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
