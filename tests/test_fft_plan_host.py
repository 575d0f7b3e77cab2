"""The device FFT engine's arithmetic, checked on the CPU.

wire-cell-toolkit_amd/csrc/fft_lds.hpp is compiled for the host (its
device functions made host-callable) into a small program that
  - runs every register butterfly Dft<R> (R = 2, 3, 4, 5, 8, 10, 15, 20 and
    the generic 7, 11, 13) against a direct DFT in both directions;
  - runs full transforms through the radix schedule sf_radix (both its
    orders: the default and the odd-first one of the unpadded row plans) with the
    Stockham pass formula of stockham_pass (y[(j - k) R + k + r Ns] =
    DFT_R(x[j + r n/R] w^(r k))) and the host twiddle-table layout, for
    lengths whose plans use the prime-factor radices 10, 15 and 20;
  - checks that the schedule's radices multiply back to n for every
    length up to 20000 with prime factors in {2, 3, 5, 7, 11, 13}, and
    that no other length gets a plan.
A dropped butterfly term, a wrong Good-Thomas index map or a plan that
loses a factor fails here before any GPU run.  The GPU tests compare the
same engine's device results with the FP64 oracle.
"""
import json
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(REPO, "wire-cell-toolkit_amd", "csrc", "fft_lds.hpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r'''
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "HDR"
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>
using namespace wct;
using cd = std::complex<double>;

static double2 tab[40];   // generic small-DFT tables 7, 11, 13 as the host builds them

template <int R> double check_dft(int sign)
{
    double2 v[R];
    std::vector<cd> x(R);
    for (int i = 0; i < R; ++i) {
        x[i] = cd(std::sin(1.3 * i + 0.2), std::cos(0.7 * i * i));
        v[i].x = x[i].real();
        v[i].y = x[i].imag();
    }
    Dft<double2, R>::run(v, sign, tab + generic_tab_offset(R));
    double err = 0;
    for (int k = 0; k < R; ++k) {
        cd acc = 0;
        for (int n = 0; n < R; ++n) acc += x[n] * std::polar(1.0, sign * 2 * M_PI * n * k / R);
        err = std::max(err, std::abs(acc - cd(v[k].x, v[k].y)));
    }
    return err;
}

template <int R> void pass(std::vector<double2>& buf, int n, int Ns, const double2* tw, int sign)
{
    const int m = n / R;
    std::vector<double2> out(n);
    for (int j = 0; j < m; ++j) {
        double2 v[R];
        for (int r = 0; r < R; ++r) v[r] = buf[j + r * m];
        const int k = j % Ns;
        if (Ns > 1)
            for (int r = 1; r < R; ++r) {
                double2 w = tw[k * (R - 1) + (r - 1)];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        Dft<double2, R>::run(v, sign, tab + generic_tab_offset(R));
        const int base = (j - k) * R + k;
        for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
    }
    buf = out;
}

double check_fft(int n, int sign, bool of = false)
{
    std::vector<double2> buf(n), tw;
    std::vector<cd> x(n);
    for (int i = 0; i < n; ++i) {
        x[i] = cd(std::sin(0.37 * i + 0.1), std::cos(0.11 * i));
        buf[i].x = x[i].real();
        buf[i].y = x[i].imag();
    }
    // per pass: w^(r k) for k < Ns, r in [1, R), w = exp(-2 pi i/(Ns R))
    int Ns = 1;
    for (int p = 0; sf_radix(n, p, of); ++p) {
        const int R = sf_radix(n, p, of);
        for (int k = 0; k < Ns; ++k)
            for (int r = 1; r < R; ++r) {
                const double ang = -2.0 * M_PI * (double)((long)r * k % ((long)Ns * R)) / (double)(Ns * R);
                double2 w;
                w.x = std::cos(ang);
                w.y = std::sin(ang);
                tw.push_back(w);
            }
        Ns *= R;
    }
    Ns = 1;
    for (int p = 0; sf_radix(n, p, of); ++p) {
        const int R = sf_radix(n, p, of);
        const double2* t = tw.data() + sf_twoff(n, p, of);
        switch (R) {
        case 2: pass<2>(buf, n, Ns, t, sign); break;
        case 3: pass<3>(buf, n, Ns, t, sign); break;
        case 4: pass<4>(buf, n, Ns, t, sign); break;
        case 5: pass<5>(buf, n, Ns, t, sign); break;
        case 7: pass<7>(buf, n, Ns, t, sign); break;
        case 8: pass<8>(buf, n, Ns, t, sign); break;
        case 10: pass<10>(buf, n, Ns, t, sign); break;
        case 11: pass<11>(buf, n, Ns, t, sign); break;
        case 13: pass<13>(buf, n, Ns, t, sign); break;
        case 15: pass<15>(buf, n, Ns, t, sign); break;
        case 20: pass<20>(buf, n, Ns, t, sign); break;
        default: return 1e9;
        }
        Ns *= R;
    }
    double err = 0, nrm = 0;
    for (int k = 0; k < n; ++k) {
        cd acc = 0;
        for (int j = 0; j < n; ++j) acc += x[j] * std::polar(1.0, sign * 2 * M_PI * (double)((long)j * k % n) / n);
        err = std::max(err, std::abs(acc - cd(buf[k].x, buf[k].y)));
        nrm = std::max(nrm, std::abs(acc));
    }
    return err / nrm;
}

int main()
{
    const int gr[3] = {7, 11, 13};
    for (int R : gr)
        for (int q = 0; q < R; ++q) {
            tab[generic_tab_offset(R) + q].x = std::cos(-2.0 * M_PI * q / R);
            tab[generic_tab_offset(R) + q].y = std::sin(-2.0 * M_PI * q / R);
        }
    printf("{\"dft\": {");
    for (int s : {-1, 1})
        printf("\"%d\": [%.3e, %.3e, %.3e, %.3e, %.3e, %.3e, %.3e, %.3e, %.3e, %.3e, %.3e]%s", s, check_dft<2>(s),
               check_dft<3>(s), check_dft<4>(s), check_dft<5>(s), check_dft<7>(s), check_dft<8>(s), check_dft<10>(s),
               check_dft<11>(s), check_dft<13>(s), check_dft<15>(s), check_dft<20>(s), s < 0 ? ", " : "");
    printf("}, \"fft\": {");
    const int ns[] = {10, 20, 40, 100, 250, 480, 800, 960, 1000, 1200, 3000, 4000, 5000, 1001, 8008, 2 * 15 * 20 * 7};
    bool first = true;
    for (int n : ns) {
        printf("%s\"%d\": [%.3e, %.3e, %.3e, %.3e]", first ? "" : ", ", n, check_fft(n, -1), check_fft(n, 1),
               check_fft(n, -1, true), check_fft(n, 1, true));
        first = false;
    }
    for (int of = 0; of < 2; ++of) {
        printf(of ? "}, \"plans_odd_first\": {" : "}, \"plans\": {");
        first = true;
        for (int n : {3000, 800, 960, 4000, 5000, 2048}) {
            printf("%s\"%d\": [", first ? "" : ", ", n);
            for (int p = 0; sf_radix(n, p, of); ++p) printf("%s%d", p ? ", " : "", sf_radix(n, p, of));
            printf("]");
            first = false;
        }
    }
    int bad = 0, maxp = 0;
    for (int n = 2; n <= 20000; ++n) {
        int m = n;
        for (int q : {2, 3, 5, 7, 11, 13})
            while (m % q == 0) m /= q;
        for (int of = 0; of < 2; ++of) {
            long prod = 1;
            int p = 0;
            for (; sf_radix(n, p, of); ++p) prod *= sf_radix(n, p, of);
            maxp = p > maxp ? p : maxp;
            if (m == 1 ? prod != n : p != 0) ++bad;
            if (of && m == 1 && sf_npass(n, true) != sf_npass(n)) ++bad;
        }
    }
    printf("}, \"bad_lengths\": %d, \"max_passes\": %d}\n", bad, maxp);
    return 0;
}
'''


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("fftplan")
    src = d / "fft_host.hip"
    src.write_text(SRC.replace("HDR", HDR))
    exe = d / "fft_host"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", "--cuda-host-only", "-o", str(exe),
                    str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def test_butterflies_match_direct_dft(report):
    for sign, errs in report["dft"].items():
        assert max(errs) < 1e-13, (sign, errs)


def test_stockham_plans_match_direct_dft(report):
    for n, errs in report["fft"].items():
        assert max(errs) < 1e-13, (n, errs)      # both schedules, both directions


def test_radix_schedule(report):
    assert report["bad_lengths"] == 0
    assert report["max_passes"] <= 24   # kMaxPasses
    plans = report["plans"]
    assert plans["3000"] == [20, 10, 15]
    assert plans["800"] == [8, 10, 10]
    assert plans["960"] == [8, 8, 15]
    assert plans["4000"] == [20, 20, 10]
    assert plans["5000"] == [10, 10, 10, 5]
    assert plans["2048"] == [8, 8, 8, 4]
    odd = report["plans_odd_first"]     # the same radices, odd ones first
    assert odd["3000"] == [15, 20, 10]
    assert odd["960"] == [15, 8, 8]
    assert odd["5000"] == [5, 10, 10, 10]
    assert odd["800"] == plans["800"] and odd["2048"] == plans["2048"]
