"""Margins of the fast kernel's prefilter (pf_terms and pf_pass in csrc/hbt_fast.h, the
constants of derive and prefilter_margins in csrc/hbt_internal.h), checked
on the host against the reference's FP64 cascade (cascade_f64, the same
function the kernels' exact stage and the oracle follow,
HBT_correlation.cpp:311-398 / :574-647).

The prefilter may pass pairs the cascade rejects, never the reverse: a pair
it drops is never binned.  Pairs are built at the windows' edges (K_T at
its ends, q_out / q_side / q_long at lo_cut and hi_cut scaled by
1 - 10^-k, k = 4..13, and inside), with random K direction, rapidity and
partner rotation, in three analyses (the bench grid; an asymmetric q window
without the LCMS boost; a narrow window with wide K_T edges), at momentum
scales up to |pT| ~ 10 GeV.  For every pair the cascade bins (same-event,
and mixed with the partner rotated as the reference does), the prefilter
on the particles' FP32 records (the precision form) and FP64 records must
pass, in both orders of the pair, and the two |p_T| must differ by at most
pt_window (the mixed items' partner window, H with margins).

The program is compiled for the host only with hipcc.  Run with the
data-scaled margins removed and the |p_T| window narrowed by 1e-3 (the
planted mistakes), the FP32 prefilter must drop binnable pairs and binnable
pairs must fall outside the window; run with the beta intervals narrowed
by 0.2 %, both prefilters must drop binnable pairs: the checks have teeth."""
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "hadronic_afterburner_toolkit_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "hbt_internal.h"
using namespace hbt;

struct P4 { double px, py, pz, E; };

template <typename Rec, typename HI, typename Prec>
static bool passes(const DevParams &P, const Rec &a, HI ahi, const Rec &b, HI bhi, double c,
                   double s, bool mixed) {
    Prec ix = a.px, iy = a.py;
    if (mixed) {  // the kernel rotates i by -theta
        ix = (Prec)((double)a.px * c + (double)a.py * s);
        iy = (Prec)((double)a.py * c - (double)a.px * s);
    }
    return pf_pass(pf_terms<Prec>(ix, iy, a.nrm, a.blo, ahi, b.px, b.py, b.nrm, b.blo, bhi,
                                  pf_const<Prec>(P)));
}

int main(int argc, char **argv) {
    const bool plant = argc > 1 && std::strcmp(argv[1], "plant") == 0;
    // the second planted mistake: beta intervals narrowed by 0.2 %
    const bool plant_beta = argc > 1 && std::strcmp(argv[1], "plantbeta") == 0;
    auto narrow = [&](auto &r, auto &hi) {
        if (!plant_beta) return;
        const auto w = (hi - r.blo) * 1e-3f;
        r.blo += w;
        hi -= w;
    };
    std::mt19937_64 rng(31);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    struct Cfg { int qnpts; double qmin, qmax, ktmin, ktmax; int boost; };
    const Cfg cfgs[] = {{41, -0.15, 0.15, 0.2, 0.7, 1}, {31, -0.05, 0.2, 0.3, 0.9, 0},
                        {11, -0.02, 0.02, 0.05, 2.0, 1}};
    const double m = 0.13957;
    long long binned = 0, dropped32 = 0, dropped64 = 0, outside_pt = 0;
    for (const Cfg &cf : cfgs) {
        hbt_config k{};
        k.qnpts = cf.qnpts; k.q_min = cf.qmin; k.q_max = cf.qmax;
        k.n_KT = 6; k.KT_min = cf.ktmin; k.KT_max = cf.ktmax;
        k.n_Kphi = 1; k.long_comoving_boost = cf.boost;
        k.HBTrap_min = -1; k.HBTrap_max = 1; k.needed_number_of_pairs = 1e18;
        k.number_of_mixed_events = 1; k.precision = 1; k.kernel = 1;
        Derived d;
        std::string err;
        if (derive(k, d, err)) { std::printf("derive: %s\n", err.c_str()); return 2; }
        // the planted mistake also narrows the |p_T| window by 1e-3
        const double ptw = pt_window(d) * (plant ? 1.0 - 1e-3 : 1.0);
        for (double scale : {1.0, 3.0, 10.0}) {
            // pairs at the edges: p1 = K + q/2, p2 = K - q/2 in the transverse
            // plane (q_out along K), K_z and q_z set from rapidity and q_long
            std::vector<std::pair<P4, P4>> prs;
            const double lo = d.P.q_lo_cut, hi = d.P.q_hi_cut;
            auto edge = [&](double v) {
                const int kk = 4 + (int)(rng() % 10);
                return v * (1.0 - std::pow(10.0, -kk));
            };
            auto qv = [&]() {
                const int w = (int)(rng() % 4);
                return w == 0 ? edge(lo) : w == 1 ? edge(hi) : lo + (hi - lo) * U(rng);
            };
            for (int n = 0; n < 40000; n++) {
                const int kw = (int)(rng() % 3);
                double kt = kw == 0 ? cf.ktmin * (1 + 1e-12 * U(rng))
                          : kw == 1 ? cf.ktmax * (1 - 1e-12 * U(rng))
                                    : cf.ktmin + (cf.ktmax - cf.ktmin) * U(rng);
                kt *= (n % 3 == 0) ? 1.0 : scale;  // larger |pT| (pairs then fall outside K_T)
                const double phi = 2 * M_PI * U(rng);
                const double qo = qv(), qs = qv(), qz = qv();
                const double cx = std::cos(phi), sx = std::sin(phi);
                const double Kx = kt * cx, Ky = kt * sx;
                const double qx = qo * cx - qs * sx, qy = qo * sx + qs * cx;
                const double Kz = (U(rng) - 0.5) * 2.0 * scale;
                P4 a{Kx + qx / 2, Ky + qy / 2, Kz + qz / 2, 0}, b{Kx - qx / 2, Ky - qy / 2, Kz - qz / 2, 0};
                a.E = std::sqrt(m * m + a.px * a.px + a.py * a.py + a.pz * a.pz);
                b.E = std::sqrt(m * m + b.px * b.px + b.py * b.py + b.pz * b.pz);
                prs.push_back({a, b});
            }
            double pmax2 = 0, bm = 0;
            for (auto &pr : prs)
                for (const P4 *p : {&pr.first, &pr.second}) {
                    pmax2 = std::max(pmax2, p->px * p->px + p->py * p->py);
                    bm = std::max(bm, std::fabs(cf.boost ? p->pz / p->E : p->pz));
                }
            DevParams P = d.P;
            prefilter_margins(d, pmax2, P);
            if (plant) {  // without the data-scaled margins
                P.f_negc = 0.0f;
                P.pf_negc = 0.0;
            }
            for (auto &pr : prs) {
                const P4 &a = pr.first, &b = pr.second;
                PfRec32 ra32, rb32; float ca32, cb32;
                PfRec64 ra64, rb64; double ca64, cb64;
                pf_record(d, a.px, a.py, a.pz, a.E, bm, ra32, ca32);
                pf_record(d, b.px, b.py, b.pz, b.E, bm, rb32, cb32);
                pf_record(d, a.px, a.py, a.pz, a.E, bm, ra64, ca64);
                pf_record(d, b.px, b.py, b.pz, b.E, bm, rb64, cb64);
                narrow(ra32, ca32);
                narrow(rb32, cb32);
                narrow(ra64, ca64);
                narrow(rb64, cb64);
                PairBins bins;
                double qinv;
                int ok;
                // same-event
                const double pta = std::sqrt(a.px * a.px + a.py * a.py);
                if (cascade_f64<false, false>(P, a.px, a.py, a.pz, a.E, b.px, b.py, b.pz, b.E, 0.0,
                                              bins, qinv, ok) & ST_BIN) {
                    binned++;
                    outside_pt += std::fabs(pta - std::sqrt(b.px * b.px + b.py * b.py)) > ptw;
                    for (int o = 0; o < 2; o++) {
                        const bool p32 = o ? passes<PfRec32, float, float>(P, rb32, cb32, ra32, ca32, 1, 0, false)
                                           : passes<PfRec32, float, float>(P, ra32, ca32, rb32, cb32, 1, 0, false);
                        const bool p64 = o ? passes<PfRec64, double, double>(P, rb64, cb64, ra64, ca64, 1, 0, false)
                                           : passes<PfRec64, double, double>(P, ra64, ca64, rb64, cb64, 1, 0, false);
                        dropped32 += !p32;
                        dropped64 += !p64;
                    }
                }
                // mixed: the partner b is the reference's rotated particle, so
                // the unrotated partner is b rotated back by -theta
                const double th = 2 * M_PI * U(rng), c = std::cos(th), s = std::sin(th);
                const double ux = b.px * c + b.py * s, uy = b.py * c - b.px * s;
                const double rx = ux * c - uy * s, ry = ux * s + uy * c;  // as the reference rotates
                if (cascade_f64<true, false>(P, a.px, a.py, a.pz, a.E, rx, ry, b.pz, b.E, 0.0,
                                             bins, qinv, ok) & ST_BIN) {
                    binned++;
                    outside_pt += std::fabs(pta - std::sqrt(ux * ux + uy * uy)) > ptw;
                    PfRec32 ru32; float cu32;
                    PfRec64 ru64; double cu64;
                    pf_record(d, ux, uy, b.pz, b.E, bm, ru32, cu32);
                    pf_record(d, ux, uy, b.pz, b.E, bm, ru64, cu64);
                    narrow(ru32, cu32);
                    narrow(ru64, cu64);
                    dropped32 += !passes<PfRec32, float, float>(P, ra32, ca32, ru32, cu32, c, s, true);
                    dropped64 += !passes<PfRec64, double, double>(P, ra64, ca64, ru64, cu64, c, s, true);
                }
            }
        }
    }
    std::printf("binned %lld dropped32 %lld dropped64 %lld outside_pt %lld\n", binned, dropped32,
                dropped64, outside_pt);
    return 0;
}
"""


def build(tmp_path):
    src = tmp_path / "pm.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "pm"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--cuda-host-only", "-ffp-contract=off",
                    f"-I{CSRC}", str(src), "-o", str(exe)], check=True, timeout=300,
                   capture_output=True)
    return exe


def run(exe, *args):
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    return int(f[1]), int(f[3]), int(f[5]), int(f[7])


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="needs hipcc")
def test_prefilter_never_drops_a_binnable_pair(tmp_path):
    exe = build(tmp_path)
    binned, d32, d64, outside_pt = run(exe)
    assert binned > 100000
    assert d32 == 0 and d64 == 0
    # every binned pair lies inside the mixed items' |p_T| window
    assert outside_pt == 0
    # the planted mistake: no data-scaled margin
    binned_p, d32_p, _, outside_p = run(exe, "plant")
    assert binned_p == binned and d32_p > 0 and outside_p > 0
    # the beta intervals narrowed by 0.2 %: binnable pairs at the q_long edge are dropped
    binned_b, d32_b, d64_b, _ = run(exe, "plantbeta")
    assert binned_b == binned and d32_b > 0 and d64_b > 0
