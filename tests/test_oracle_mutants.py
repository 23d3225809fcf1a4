"""The oracle's pins must catch plausible mistakes.

Each case below plants one slip (a sign, a swapped operand, a dropped
term, an off-by-one index, a strict-versus-inclusive comparison) in a copy
of oracle/oracle.c, builds it into a temporary directory, and runs the pin
and golden suites against that build in a child process.  A mutant that
passes every test would mean the pins do not constrain that part of the
algorithm.  CPU only: the mistakes are planted in the C oracle, never in
GPU code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (text in oracle.c, replacement): each text occurs exactly once
MUTANTS = {
    "srotation_sign": ("double const y_hat = -sin_z * x + cos_z * y;",
                       "double const y_hat = sin_z * x + cos_z * y;"),
    "xyshift_sign": ("p->x[ i ] += minus_dx;", "p->x[ i ] -= minus_dx;"),
    "dipedge_operand": ("p->py[ i ] += e->r43 * p->y[ i ];",
                        "p->py[ i ] += e->r43 * p->x[ i ];"),
    "multipole_horner": ("double const zim = dpx * y + dpy * x;",
                         "double const zim = dpx * y - dpy * x;"),
    "curvature_dispersion": ("dpx += hxl + hxl * delta;", "dpx += hxl;"),
    "curvature_path_length": ("p->zeta[ i ] += chi * ( hyly - hxlx );",
                              "p->zeta[ i ] += chi * ( hxlx - hyly );"),
    "curvature_focusing": ("dpx -= b1l * hxlx / length;", "dpx -= b1l * hxlx;"),
    "cavity_phase": ("double const energy = q * c->voltage * OR_SIN( phase );",
                     "double const energy = q * c->voltage * OR_COS( phase );"),
    "energy_rvv": ("double const rvv = one_plus_delta / ( 1.0 + ptau_beta0 );\n"
                   "    p->delta[ i ] = delta;",
                   "double const rvv = one_plus_delta / ( 1.0 + ptau );\n"
                   "    p->delta[ i ] = delta;"),
    "rect_strict": ("( x >= l->min_x ) &&", "( x > l->min_x ) &&"),
    "ellipse_axes": ("temp *= l->b_squ;", "temp *= l->a_squ;"),
    "global_aperture_inclusive": ("( ( sign_x * x ) < 1.0 ) &", "( ( sign_x * x ) <= 1.0 ) &"),
    "sc_lorentz_factor": ("charge * charge * ( 1.0 - beta0 * beta );\n"
                          "    fact_kick *= sc->number_of_particles;",
                          "charge * charge * ( 1.0 - beta * beta );\n"
                          "    fact_kick *= sc->number_of_particles;"),
    "interpolation_node": ("result = y[ idx ] + ( z - ( d->z0 + d->dz * idx ) ) * yd[ idx ];",
                           "result = y[ idx ] + ( z - ( d->z0 + d->dz * idx ) ) * yd[ idx + 1 ];"),
    "ebe_index_order": ("out = turn_off * np * ne + elem_off * np + pid_off;",
                        "out = turn_off * np * ne + pid_off * ne + elem_off;"),
    "rf_multipole_sign": ("double const py_kick =  chi * dpy;",
                          "double const py_kick = -chi * dpy;"),
    "rf_energy_cos_for_sin": ("dptr += sn * ( bal_n_kk * zre ) - ss * ( bal_s_kk * zim );",
                              "dptr += cn * ( bal_n_kk * zre ) - ss * ( bal_s_kk * zim );"),
    "rf_energy_sign": ("double const rfmultipole_energy = -q * ( ( k * p0c ) * dptr );",
                       "double const rfmultipole_energy = q * ( ( k * p0c ) * dptr );"),
    "rf_skew_phase_index": ("double const ps_kk = DEG2RAD * phase[ 2 * kk + 1 ] - ktau;",
                            "double const ps_kk = DEG2RAD * phase[ 2 * kk ] - ktau;"),
    "qgauss_exponent": ("e = OR_POW( u_plus, 1.0 / one_minus_q );",
                        "e = OR_POW( u_plus, one_minus_q );"),
    "qgauss_support": ("if( u_plus < 0.0 ) u_plus = 0.0;", ""),
    "monitor_rolling_slot": ("store_idx = store_idx % m->num_stores;",
                             "store_idx = ( store_idx + 1 ) % m->num_stores;"),
    "monitor_particle_order": ("else                     out = pid_off * nstore + store_idx;",
                               "else                     out = pid_off * m->num_stores + store_idx;"),
    "ebe_rolling_wrap": ("out = out % capacity;", "out = ( out + np ) % capacity;"),
}


def _build(src, outdir):
    path = os.path.join(outdir, "oracle.c")
    with open(path, "w") as f:
        f.write(src)
    for lib, extra in (("liboracle.so", []), ("liboracle_libm.so", ["-DORACLE_USE_LIBM"])):
        subprocess.check_call(["gcc", "-std=c99", "-O2", "-fPIC", "-fno-fast-math",
                               "-ffp-contract=off", "-fopenmp", "-I", os.path.join(ROOT, "oracle"),
                               "-shared", "-o", os.path.join(outdir, lib), path, "-lm"] + extra)


def _run_suite(oracle_dir):
    code = ("import sys; sys.path.insert(0, 'tests'); import oracle_lib; "
            "oracle_lib.ORACLE_DIR = %r; import pytest; "
            "sys.exit(pytest.main(['-q', '-x', '-p', 'no:cacheprovider', "
            "'tests/test_oracle_pins.py', 'tests/test_oracle_golden.py']))" % oracle_dir)
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    return r.returncode, r.stdout[-2000:]


def test_pins_catch_planted_mistakes(tmp_path):
    with open(os.path.join(ROOT, "oracle", "oracle.c")) as f:
        src = f.read().replace('"../include/', '"%s/include/' % ROOT)
    dirs = {}
    for name, (old, new) in MUTANTS.items():
        assert src.count(old) == 1, name
        d = tmp_path / name
        d.mkdir()
        _build(src.replace(old, new), str(d))
        dirs[name] = str(d)
    # the unmutated copy passes: a failure below is the mutant's doing
    base = tmp_path / "unmutated"
    base.mkdir()
    _build(src, str(base))
    rc, out = _run_suite(str(base))
    assert rc == 0, out
    with ThreadPoolExecutor(max_workers=4) as pool:
        results = dict(zip(dirs, pool.map(_run_suite, dirs.values())))
    missed = [name for name, (rc, _) in results.items() if rc != 1]
    assert not missed, {name: results[name][1] for name in missed}
