#include "coupling.hpp"

#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

#include "core.hpp"

namespace mq {

namespace {

// n! for n <= 170 in long double
long double factorial(int n) {
    static std::vector<long double> table = [] {
        std::vector<long double> t(171);
        t[0] = 1.0L;
        for (int i = 1; i <= 170; ++i) t[i] = t[i - 1] * i;
        return t;
    }();
    if (n < 0 || n > 170) throw std::out_of_range("factorial argument out of range");
    return table[n];
}

bool triangle_fails(int a, int b, int c) { return b < std::abs(a - c) || b > a + c || ((a + b + c) & 1); }

// Delta(abc) with integer-valued j's given as 2j
long double delta(int ta, int tb, int tc) {
    return std::sqrt(factorial((ta + tb - tc) / 2) * factorial((ta - tb + tc) / 2) *
                     factorial((-ta + tb + tc) / 2) / factorial((ta + tb + tc) / 2 + 1));
}

long double sixj_racah(int a, int b, int c, int d, int e, int f) {
    // triads (a b c) (a e f) (d b f) (d e c), all in 2j units
    int s1 = (a + b + c) / 2, s2 = (a + e + f) / 2, s3 = (d + b + f) / 2, s4 = (d + e + c) / 2;
    int p1 = (a + b + d + e) / 2, p2 = (b + c + e + f) / 2, p3 = (c + a + f + d) / 2;
    int tmin = std::max(std::max(s1, s2), std::max(s3, s4));
    int tmax = std::min(p1, std::min(p2, p3));
    long double sum = 0.0L;
    for (int t = tmin; t <= tmax; ++t) {
        long double term = factorial(t + 1) /
                           (factorial(t - s1) * factorial(t - s2) * factorial(t - s3) * factorial(t - s4) *
                            factorial(p1 - t) * factorial(p2 - t) * factorial(p3 - t));
        sum += (t & 1) ? -term : term;
    }
    return delta(a, b, c) * delta(a, e, f) * delta(d, b, f) * delta(d, e, c) * sum;
}

}  // namespace

double wigner_6j(int a, int b, int c, int d, int e, int f) {
    if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0) return std::numeric_limits<double>::quiet_NaN();
    if (triangle_fails(a, b, c) || triangle_fails(a, e, f) || triangle_fails(d, b, f) || triangle_fails(d, e, c))
        return 0.0;
    return double(sixj_racah(a, b, c, d, e, f));
}

double wigner_9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0 || g < 0 || h < 0 || i < 0)
        return std::numeric_limits<double>::quiet_NaN();
    if (triangle_fails(a, b, c) || triangle_fails(d, e, f) || triangle_fails(g, h, i) ||
        triangle_fails(a, d, g) || triangle_fails(b, e, h) || triangle_fails(c, f, i))
        return 0.0;
    // {a b c; d e f; g h i} = sum_x (-1)^{2x} (2x+1) {a b c; f i x}{d e f; b x h}{g h i; x a d}
    int xmin = std::max(std::abs(a - i), std::max(std::abs(d - h), std::abs(b - f)));
    int xmax = std::min(a + i, std::min(d + h, b + f));
    long double sum = 0.0L;
    for (int x = xmin; x <= xmax; x += 2) {
        long double t1 = (triangle_fails(a, b, c) || triangle_fails(a, i, x) || triangle_fails(f, b, x) ||
                          triangle_fails(f, i, c)) ? 0.0L : sixj_racah(a, b, c, f, i, x);
        long double t2 = (triangle_fails(d, e, f) || triangle_fails(d, x, h) || triangle_fails(b, e, h) ||
                          triangle_fails(b, x, f)) ? 0.0L : sixj_racah(d, e, f, b, x, h);
        long double t3 = (triangle_fails(g, h, i) || triangle_fails(g, a, d) || triangle_fails(x, h, d) ||
                          triangle_fails(x, a, i)) ? 0.0L : sixj_racah(g, h, i, x, a, d);
        long double term = (x + 1) * t1 * t2 * t3;
        sum += (x & 1) ? -term : term;
    }
    return double(sum);
}

double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * wigner_9j(a, b, c, d, e, f, g, h, i);
}

void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    double prefactor = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    if (triangle(a, b, c)) {
        out[0] = prefactor * mod_coupling(a, b, c, d, e, f, g, h, i);
        out[2] = prefactor * mod_coupling(a, b, c, d, e, f, g, 2, i);
    }
    if (triangle(a, 2, c)) {
        out[1] = prefactor * mod_coupling(a, 2, c, d, e, f, g, h, i);
        out[3] = prefactor * mod_coupling(a, 2, c, d, e, f, g, 2, i);
    }
}

Wigner9jCache::Wigner9jCache(int J, int I, int Ip) {
    int two_sp = std::abs(I - Ip);
    jp_max_ = J + 2;
    jp_min_ = std::max(J - 2, 0);
    int nj = jp_max_ - jp_min_ + 1;
    coeff_.assign(size_t(48) * nj, 0.0);
    // (A, K, A') combinations of the table, A = i / 4
    static const int Ks[12] = {0, 1, 2, 0, 0, 1, 2, 1, 0, 1, 2, 0};
    static const int Aps[12] = {0, 1, 2, 0, 1, 0, 1, 2, 2, 1, 0, 0};
    for (int Jp = jp_min_; Jp <= jp_max_; ++Jp) {
        int two_s = std::abs(J - Jp);
        for (int t = 0; t < 12; ++t)
            set_coupling(J, two_s, Jp, t / 4, Ks[t], Aps[t], I, two_sp, Ip, 1.0,
                         &coeff_[4 * t + 48 * (Jp - jp_min_)]);
    }
}

std::shared_ptr<const Wigner9jCache> wigner9j_cache(int J, int I, int Ip) {
    static std::mutex m;
    static std::map<std::tuple<int, int, int>, std::shared_ptr<const Wigner9jCache>> tables;
    const auto key = std::make_tuple(J, I, Ip);
    {
        std::lock_guard<std::mutex> lock(m);
        auto it = tables.find(key);
        if (it != tables.end()) return it->second;
    }
    auto t = std::make_shared<const Wigner9jCache>(J, I, Ip);  // built outside the lock
    std::lock_guard<std::mutex> lock(m);
    return tables.emplace(key, t).first->second;
}

void Wigner9jCache::set_scale(int A, int K, int Ap, int Jp, double scale, double out[4]) const {
    require(Jp >= jp_min_ && Jp <= jp_max_, "Wigner9jCache: J' outside J +- 2");
    int h = hash(A, K, Ap);
    require(h >= 0 && h < 12, "Wigner9jCache: (A, K, A') outside the coupling table");
    const double* c = &coeff_[4 * h + 48 * (Jp - jp_min_)];
    for (int k = 0; k < 4; ++k) out[k] = c[k] * scale;
}

double conjugate_correction(Sym sym, Charge lc, Charge rc) {
    if (sym != Sym::SU2U1) return 1.0;
    int sl = spin_of(lc), sr = spin_of(rc);
    require(sl >= 0 && sr >= 0, "conjugate_correction: negative spin");
    int S = std::min(sl, sr);
    int d = sr - sl;
    switch (d) {
        case 0: return 1.0;
        case 1: return -std::sqrt((S + 1.) / (S + 2.));
        case -1: return std::sqrt((S + 2.) / (S + 1.));
        case 2: return -std::sqrt((S + 1.) / (S + 3.));
        case -2: return -std::sqrt((S + 3.) / (S + 1.));
        default:
            throw Error(MQ_ERR_INVALID, "hermitian conjugate for reduced tensor operators only up to rank 1");
    }
}

}  // namespace mq
