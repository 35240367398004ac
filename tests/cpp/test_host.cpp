// CPU checks of the C++ front end (no GPU calls): the command line of
// main.cpp:111-136 and the Distribution generator of Distribution.cpp:55-130.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../csrst_amd/host/csrst.hpp"

static int fails = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++fails; } \
    } while (0)

int main()
{
    {   // defaults of main.cpp:111-136
        char a0[] = "csrst";
        char *argv[] = {a0};
        const csrst::Commands c = csrst::Commands::parse(1, argv);
        CHECK(c.numsteps == 10 && c.numgridpx == 100 && c.numgridpy == 100 && c.statmodulus == 100000);
        CHECK(c.phi == 0.0 && c.width == 0.10 && c.length == 0.10 && c.lengthbefore == 0.10);
        CHECK(c.Ekin == 250.0 && c.Bz == 0.10 && c.Qtotal == -1.0 && c.Np == 50000 && !c.no_initials);
        CHECK(c.mean_Rx == 0.05 && c.sigma_x == 0.01 && c.sigma_p == 1.0 && c.order == 0 && c.s2bug);
    }
    {   // short, long and = forms; switches
        const char *args[] = {"csrst", "-x", "512", "--numgridpy=256", "-P", "10", "--lengthbefore", "0.05",
                              "-I", "--order", "2", "--no-s2bug", "-E", "50.59", "--Np", "1000"};
        const csrst::Commands c = csrst::Commands::parse(16, const_cast<char **>(args));
        CHECK(c.numgridpx == 512 && c.numgridpy == 256 && c.phi == 10.0 && c.lengthbefore == 0.05);
        CHECK(c.no_initials && c.order == 2 && !c.s2bug && c.Ekin == 50.59 && c.Np == 1000);
        const csrst_params p = c.params();
        CHECK(p.Nx == 512 && p.Ny == 256 && p.order == 2 && p.s2bug == 0 && p.ekinMeV == 50.59);
    }
    {   // errors
        const char *bad1[] = {"csrst", "--bogus"};
        const char *bad2[] = {"csrst", "-x", "12a"};
        const char *bad3[] = {"csrst", "-x"};
        const char *bad4[] = {"csrst", "--order", "3"};
        int thrown = 0;
        for (auto *a : {bad1, bad2, bad4}) {
            try { csrst::Commands::parse(a == bad1 ? 2 : 3, const_cast<char **>(a)); } catch (const csrst::Error &e) {
                thrown += e.code == CSRST_EINVAL;
            }
        }
        try { csrst::Commands::parse(2, const_cast<char **>(bad3)); } catch (const csrst::Error &) { ++thrown; }
        CHECK(thrown == 4);
    }
    {   // Halton coordinates (radical inverse) of the first indices
        CHECK(csrst::Distribution::radical_inverse(1, 2) == 0.5);
        CHECK(csrst::Distribution::radical_inverse(2, 2) == 0.25);
        CHECK(csrst::Distribution::radical_inverse(3, 2) == 0.75);
        CHECK(std::fabs(csrst::Distribution::radical_inverse(1, 3) - 1.0 / 3) < 1e-16);
        CHECK(std::fabs(csrst::Distribution::radical_inverse(5, 3) - 7.0 / 9) < 1e-15);   // 5 = 12_3 -> 0.21_3
    }
    {   // van der Corput: the first b^m - 1 radical inverses in base b are the
        // fractions j / b^m, j = 1 .. b^m - 1, each once (digit reversal is a
        // bijection); an index offset, a dropped or reordered digit, or a base
        // mix-up breaks it.  Exact in binary for base 2; to 1e-15 for base 3.
        for (unsigned b : {2u, 3u}) {
            const unsigned long long N = b == 2 ? 4096 : 2187;   // 2^12, 3^7
            std::vector<int> seen(N, 0);
            bool ok = true;
            for (unsigned long long k = 1; k < N; ++k) {
                const double v = csrst::Distribution::radical_inverse(k, b) * (double)N;
                const long long j = std::llround(v);
                ok = ok && std::fabs(v - (double)j) < 1e-9 && j >= 1 && j < (long long)N && !seen[j];
                if (j >= 1 && j < (long long)N) seen[j] = 1;
            }
            CHECK(ok);
        }
        // the digit-reversal formula itself, for indices with many digits
        for (unsigned long long k : {12345ull, 987654321ull, (1ull << 40) + 12345ull}) {
            for (unsigned b : {2u, 3u}) {
                unsigned long long kk = k, num = 0, den = 1;
                while (kk) { num = num * b + kk % b; den *= b; kk /= b; }
                const double want = (double)num / (double)den;
                CHECK(std::fabs(csrst::Distribution::radical_inverse(k, b) - want) <= 4e-16);
            }
        }
    }
    {   // generated bunch: moments, charge, momentum offset and the XFirst order of R
        csrst::Commands c;
        c.Np = 20000; c.mean_Rx = 0.02; c.mean_Ry = 0.006; c.sigma_x = 1e-4; c.sigma_p = 0.5; c.Ekin = 50.0;
        c.Qtotal = 0.5;
        const double hr[2] = {2e-5, 2.5e-5}, org[2] = {-2e-5, -2.5e-5};
        const csrst::Bunch b = csrst::Distribution::generate(c, hr, org);
        double mx = 0, my = 0, sx = 0, px = 0, q = 0;
        for (size_t k = 0; k < b.size(); ++k) {
            mx += b.R[2 * k]; my += b.R[2 * k + 1]; px += b.P[2 * k]; q += b.Q[k];
        }
        mx /= b.size(); my /= b.size(); px /= b.size();
        for (size_t k = 0; k < b.size(); ++k) sx += (b.R[2 * k] - mx) * (b.R[2 * k] - mx);
        sx = std::sqrt(sx / b.size());
        const double gamma = 50.0 / (0.51099906e-03 * 1e3) + 1;
        CHECK(std::fabs(mx - 0.02) < 2e-6 && std::fabs(my - 0.006) < 2e-6);
        CHECK(std::fabs(sx / 1e-4 - 1) < 0.02);
        CHECK(std::fabs(px - std::sqrt(gamma * gamma - 1)) < 0.01);
        CHECK(std::fabs(q + 0.5e-9) < 1e-20);
        bool sorted = true;
        for (size_t k = 1; k < b.size(); ++k) {
            const int J0 = (int)((b.R[2 * k - 1] - org[1]) / hr[1]), J1 = (int)((b.R[2 * k + 1] - org[1]) / hr[1]);
            if (J1 < J0 || (J1 == J0 && b.R[2 * k] < b.R[2 * k - 2])) sorted = false;
        }
        CHECK(sorted);
    }
    std::printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
    return fails ? 1 : 0;
}
