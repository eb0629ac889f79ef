"""Builds copies of the oracle with one planted mistake, on the CPU, so that a pin
test can show that it fails on the mistake.  Each mutant is the oracle source with
exact text replacements (every one must match the stated number of times), built
with g++ into a temporary directory; nothing here touches the built oracle."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cmdg_oracle.cpp")
_built = {}

# name -> [(old text, new text, occurrences)]
MUTANTS = {
    # constant-viscosity stress, TurbulenceClosures.jl:366-404
    "divergence_factor_dropped": [
        ("v = v + ((T(2) * nu) / T(3)) * trS;", "v = v + (T(2) * nu) * trS;", 1)],
    "divergence_sign_flipped": [
        ("v = v + ((T(2) * nu) / T(3)) * trS;", "v = v - ((T(2) * nu) / T(3)) * trS;", 1)],
    "dynamic_viscosity_not_divided_by_rho": [
        ("T nu = (c.d.turbulence == CMDG_CONSTANT_DYNAMIC_VISCOSITY) ? c.visc / q[0] : c.visc;",
         "T nu = c.visc;", 1)],
    # N^2 = grad theta_v . grad Phi / theta_v, TurbulenceClosures.jl:455-470
    "buoyancy_with_unit_vertical": [
        ("gf[9] = ((gG[0][4] * a.gPhi[0] + gG[1][4] * a.gPhi[1]) + gG[2][4] * a.gPhi[2]) / a.theta_v;",
         "gf[9] = ((gG[0][4] * (a.gPhi[0] / c.grav) + gG[1][4] * (a.gPhi[1] / c.grav)) + "
         "gG[2][4] * (a.gPhi[2] / c.grav)) / a.theta_v;", 1)],
    # metric terms xi_a x_d (vgeo field a + 3 d) used transposed
    "tendency_metric_transposed": [
        ("T x1 = vg(c, n, ax, e), x2 = vg(c, n, ax + 3, e), x3 = vg(c, n, ax + 6, e);",
         "T x1 = vg(c, n, 3 * ax, e), x2 = vg(c, n, 3 * ax + 1, e), x3 = vg(c, n, 3 * ax + 2, e);", 2)],
    "gradient_metric_transposed": [
        ("v += vg(c, n, 0 + 3 * d, e) * g1;", "v += vg(c, n, d, e) * g1;", 1),
        ("v += vg(c, n, 1 + 3 * d, e) * g2;", "v += vg(c, n, 3 + d, e) * g2;", 1),
        ("gG[d][s] = -T(0) + vg(c, n, 2 + 3 * d, e) * g3;", "gG[d][s] = -T(0) + vg(c, n, 6 + d, e) * g3;", 1)],
}


def build(name: str) -> str:
    """Path of the oracle library built with mutant `name`."""
    if name in _built:
        return _built[name]
    src = open(SRC).read()
    for old, new, count in MUTANTS[name]:
        found = src.count(old)
        if found != count:
            raise RuntimeError(f"mutant {name}: expected {count} match(es) of {old!r}, found {found}")
        src = src.replace(old, new)
    d = tempfile.mkdtemp(prefix=f"oracle_mutant_{name}_")
    path_src = os.path.join(d, "cmdg_oracle.cpp")
    with open(path_src, "w") as f:
        f.write(src)
    lib = os.path.join(d, "liboracle.so")
    # the source includes "../include/cmdg.h": resolve it from the oracle directory
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                    "-iquote", os.path.join(ROOT, "oracle"), "-o", lib, path_src], check=True)
    _built[name] = lib
    return lib
