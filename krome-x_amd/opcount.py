"""Static operation counts of a generated network's right-hand side: the
algorithmic FLOPs per rhs evaluation the roofline of the benchmark uses.

The generated functions are templates on the real type, so a small host
program (g++) evaluates them once on a counting type whose arithmetic
operators and math functions increment counters: each add, subtract,
multiply, divide, comparison-free min/max and each exp / log / log10 /
pow / sqrt / tanh counts as ONE algorithmic operation (a transcendental is
one operation of the algorithm, whatever its instruction count).  The
count follows the branches the sample state takes (one piece of each
piecewise rate), at a primordial3-like state.

    python3 krome-x_amd/opcount.py   -> csrc/generated/opcount.json
"""
from __future__ import annotations

import json
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
GEN = HERE / "csrc" / "generated"

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#define KX_HD
namespace kx { struct C; }
struct Cnt { static long ops, tr; };
long Cnt::ops = 0, Cnt::tr = 0;
namespace kx {
struct C {
  double v;
  C() : v(0) {}
  C(double x) : v(x) {}
  explicit operator double() const { return v; }
};
inline C operator+(C a, C b) { Cnt::ops++; return C(a.v + b.v); }
inline C operator-(C a, C b) { Cnt::ops++; return C(a.v - b.v); }
inline C operator*(C a, C b) { Cnt::ops++; return C(a.v * b.v); }
inline C operator/(C a, C b) { Cnt::ops++; return C(a.v / b.v); }
inline C operator-(C a) { return C(-a.v); }
inline C& operator+=(C& a, C b) { a = a + b; return a; }
inline C& operator-=(C& a, C b) { a = a - b; return a; }
inline C& operator*=(C& a, C b) { a = a * b; return a; }
inline bool operator<(C a, C b) { return a.v < b.v; }
inline bool operator>(C a, C b) { return a.v > b.v; }
inline bool operator<=(C a, C b) { return a.v <= b.v; }
inline bool operator>=(C a, C b) { return a.v >= b.v; }
inline bool operator==(C a, C b) { return a.v == b.v; }
inline bool operator!=(C a, C b) { return a.v != b.v; }
#define KX_T1(f) inline C f(C x) { Cnt::ops++; Cnt::tr++; return C(std::f(x.v)); }
KX_T1(exp) KX_T1(log) KX_T1(log10) KX_T1(sqrt) KX_T1(tanh)
inline C pow(C x, C y) { Cnt::ops++; Cnt::tr++; return C(std::pow(x.v, y.v)); }
inline C fabs(C x) { return C(std::fabs(x.v)); }
inline C fmin(C x, C y) { Cnt::ops++; return C(std::fmin(x.v, y.v)); }
inline C fmax(C x, C y) { Cnt::ops++; return C(std::fmax(x.v, y.v)); }
}
#define KX_PHYSICS_COUNTING 1
#include "kx_physics.h"
#include "net_primordial3.h"
using Net = kx::net_primordial3;
int main() {
  using kx::C;
  C n[Net::NSPEC + 1], k[Net::NREA], f[Net::NSPEC + 1];
  const double nd[9] = {9e5, 1e-2, 1e-9, 5e4, 1e-12, 7.75e4, 1e-12, 1e-20, 1e-2};
  for (int i = 0; i < Net::NSPEC; ++i) n[i] = C(nd[i]);
  Net::Params p; p.Tcmb = 2.725; p.tff = 1e12; p.opratio = 3.0;
  const C T(300.0);
  long o0 = Cnt::ops, t0 = Cnt::tr;
  for (int i = 0; i < Net::NREA; ++i) k[i] = C(0.0);
  Net::rates_T(T, p, k);
  const long rT = Cnt::ops - o0, rTt = Cnt::tr - t0;
  o0 = Cnt::ops; t0 = Cnt::tr;
  Net::rates_n(T, n, p, k);
  const long rN = Cnt::ops - o0, rNt = Cnt::tr - t0;
  o0 = Cnt::ops; t0 = Cnt::tr;
  Net::rhs(n, k, f);
  const long rh = Cnt::ops - o0;
  o0 = Cnt::ops; t0 = Cnt::tr;
  f[Net::NSPEC] = Net::thermal(T, n, k, p);
  const long th = Cnt::ops - o0, tht = Cnt::tr - t0;
  std::printf("{\"rates_T\": %ld, \"rates_T_transc\": %ld, \"rates_n\": %ld, "
              "\"rates_n_transc\": %ld, \"rhs\": %ld, \"thermal\": %ld, "
              "\"thermal_transc\": %ld}\n", rT, rTt, rN, rNt, rh, th, tht);
  return 0;
}
"""


def count():
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "opcount.cpp"
        exe = Path(d) / "opcount"
        src.write_text(PROGRAM)
        subprocess.run(["g++", "-std=c++17", "-O0", "-I", str(HERE / "csrc"),
                        "-I", str(GEN), str(src), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True,
                             text=True).stdout
    c = json.loads(out)
    c["full_rhs"] = c["rates_T"] + c["rates_n"] + c["rhs"] + c["thermal"]
    return {"primordial3": c}


if __name__ == "__main__":
    res = count()
    (GEN / "opcount.json").write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res))
