// SU(2) recoupling coefficients of the reduced-tensor-operator contraction.
//
// 6j and 9j symbols are evaluated from the Racah single-sum formula and the standard 6j-sum
// expansion of the 9j symbol (Edmonds, "Angular Momentum in Quantum Mechanics", eqs. 6.3.7 and
// 6.4.3), arguments given as 2j.  Selection rules follow the reference's GSL extract
// (block_matrix/symmetry/wignerj/coupling.c:197-205, 427-446): a failed triangle or parity gives 0,
// a negative argument gives NaN.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

namespace mq {

double wigner_6j(int two_ja, int two_jb, int two_jc, int two_jd, int two_je, int two_jf);
double wigner_9j(int two_ja, int two_jb, int two_jc, int two_jd, int two_je, int two_jf,
                 int two_jg, int two_jh, int two_ji);

// sqrt((g+1)(h+1)(c+1)(f+1)) * 9j  (reference: SU2::mod_coupling, gsl_coupling.h:38-47)
double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i);

inline bool triangle(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b; }

// Four couplings selected by (row_spin == 2, col_spin == 2) of a sparse operator entry
// (reference: SU2::set_coupling, gsl_coupling.h:63-82).
void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]);

// Per (J, I, I') cache of the 12 (A, K, A') bond/operator spin combinations that occur in the
// reference's MPOs (reference: SU2::Wigner9jCacheI, gsl_coupling.h:141-204).  Lookups go through
// the reference's hash 4A + K + (A==1 && K==1 ? A' : 0), including its aliasing of combinations
// outside the table.
class Wigner9jCache {
public:
    Wigner9jCache(int J, int I, int Ip);
    void set_scale(int A, int K, int Ap, int Jp, double scale, double out[4]) const;

private:
    int jp_min_, jp_max_;
    std::vector<double> coeff_;
    static int hash(int a, int b, int c) { return 4 * a + b + ((a == 1 && b == 1) ? c : 0); }
};
// the table of (J, I, I'), made once per process and shared (thread-safe): the schedule builders
// ask for the same few spin triples at every cohort of every site
std::shared_ptr<const Wigner9jCache> wigner9j_cache(int J, int I, int Ip);

}  // namespace mq
