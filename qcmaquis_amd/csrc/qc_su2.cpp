// Spin-adapted (SU2U1) quantum-chemistry Hamiltonian MPO, the input producer of the SU2U1 path.
//
//   H = sum_pq h_pq E_pq + 1/2 sum_pqrs (pq|rs) sum_st c+_ps c+_rt c_st c_qs
//
// Site basis (descending SU2U1 order, charge = (N, 2S) added along the chain):
//   0 = (2,0) doubly occupied, 1 = (1,1) singly occupied coupled up (S + 1/2),
//   2 = (1,-1) singly occupied coupled down (S - 1/2), 3 = (0,0) empty.
// Operators are reduced tensor operators of rank K whose action moves the MPO bond spin by
// out - in; the MPO bond spin is the coupled spin of the operators applied left of the bond.
//
// Reference (paths relative to dmrg/framework/dmrg/):
//   * reduced site operators and their bond-spin actions: models/chem/su2u1/model.hpp:28-300
//   * decomposition of the spin-summed one- and two-electron terms into site-ordered strings of
//     those operators, per index pattern of (ij|kl): models/chem/su2u1/model.hpp:347-588,
//     su2u1/term_maker.h:28-463
//   * merging of equal strings: su2u1/chem_helper.h:69-106
//   * Hermitian-partner phases of MPO bond keys: generate_mpo/tagged_mpo_maker_optim.hpp:608-690
//   * two-site MPO by SU(2) Kronecker products: mp_tensors/ts_ops.hpp:54-168,
//     block_matrix/site_operator_algorithms.h:214-335
// The bond-state keying is this package's own (the one of the TwoU1 builder in qc_mpo.cpp): a
// 4-operator string whose gap between its 2nd and 3rd operator lies right of the chain centre
// carries its pending operator pair across the gap instead of the applied pair, so that no bond
// holds more than ~min(p, L-p)^2 pair states.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <set>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "core.hpp"
#include "coupling.hpp"
#include "qc_mpo.hpp"

namespace mq {

namespace {

const Charge SQ[4] = {Charge(2, 0), Charge(1, 1), Charge(1, -1), Charge(0, 0)};
enum { SA = 0, SB = 1, SC = 2, SD = 3 };

int abs_spin(int s) { return std::abs(SQ[s].q[1]); }
// coupled spin of a two-site product state: the pair (1,-1)(1,1) is the triplet, every other
// pair couples to |2S_a + 2S_b| (spin_descriptor.h:127-137)
int product_spin(int a, int b) {
    int sa = SQ[a].q[1], sb = SQ[b].q[1];
    if (sa == -1 && sb == 1) return 2;
    return std::abs(sa + sb);
}

struct SOp {
    int K = 0, action = 0;  // 2 x rank; change of the bond spin
    bool fermionic = false;
    double m[4][4] = {};    // m[in][out]
};

enum Op : int {
    IDENT, IDENT_FULL, FILL,
    CREATE_FILL, CREATE, DESTROY_FILL, DESTROY,
    CREATE_FILL_DN, DESTROY_FILL_DN, CREATE_UP, DESTROY_UP,
    CREATE_FILL_COUNT, CREATE_COUNT, DESTROY_FILL_COUNT, DESTROY_COUNT,
    COUNT, DOCC, E2D, D2E, FLIP_S0, FLIP_TO_S2, FLIP_TO_S0, COUNT_FILL,
    N_BASE_OPS
};

struct OpTable {
    std::vector<SOp> ops;
    std::vector<int> herm;       // adjoint op, -1: not registered (the adjoint of a key keeps the op)
    std::vector<char> selfadj;   // registered self-adjoint

    OpTable() {
        ops.resize(N_BASE_OPS);
        herm.assign(N_BASE_OPS, -1);
        selfadj.assign(N_BASE_OPS, 0);
        const double r2 = std::sqrt(2.0), r32 = std::sqrt(1.5);
        auto set = [&](int o, int K, int in, int out, bool f, std::initializer_list<std::tuple<int, int, double>> el) {
            SOp& s = ops[o];
            s.K = K;
            s.action = out - in;
            s.fermionic = f;
            for (auto const& e : el) s.m[std::get<0>(e)][std::get<1>(e)] = std::get<2>(e);
        };
        set(IDENT, 0, 0, 0, false, {{SA, SA, 1}, {SB, SB, 1}, {SC, SC, 1}, {SD, SD, 1}});
        set(IDENT_FULL, 0, 0, 0, false, {{SA, SA, 1}, {SB, SB, 1}, {SC, SC, 1}, {SD, SD, 1}, {SB, SC, 1}, {SC, SB, 1}});
        set(FILL, 0, 0, 0, false, {{SA, SA, 1}, {SD, SD, 1}, {SB, SB, -1}, {SC, SC, -1}, {SB, SC, -1}, {SC, SB, -1}});
        // creators / annihilators; "fill" variants carry the Jordan-Wigner sign of their own site
        set(CREATE_FILL, 1, 0, 1, true, {{SB, SA, r2}, {SC, SA, r2}, {SD, SB, 1}, {SD, SC, 1}});
        set(CREATE, 1, 1, 0, true, {{SB, SA, r2}, {SC, SA, r2}, {SD, SB, -1}, {SD, SC, -1}});
        set(DESTROY_FILL, 1, 0, 1, true, {{SA, SB, 1}, {SA, SC, 1}, {SB, SD, -r2}, {SC, SD, -r2}});
        set(DESTROY, 1, 1, 0, true, {{SA, SB, 1}, {SA, SC, 1}, {SB, SD, r2}, {SC, SD, r2}});
        ops[CREATE_FILL_DN] = ops[CREATE_FILL];
        ops[CREATE_FILL_DN].action = -1;
        ops[DESTROY_FILL_DN] = ops[DESTROY_FILL];
        ops[DESTROY_FILL_DN].action = -1;
        ops[CREATE_UP] = ops[CREATE];
        ops[CREATE_UP].action = 1;
        ops[DESTROY_UP] = ops[DESTROY];
        ops[DESTROY_UP].action = 1;
        set(CREATE_FILL_COUNT, 1, 0, 1, true, {{SB, SA, r2}, {SC, SA, r2}});
        set(CREATE_COUNT, 1, 1, 0, true, {{SB, SA, r2}, {SC, SA, r2}});
        set(DESTROY_FILL_COUNT, 1, 0, 1, true, {{SA, SB, 1}, {SA, SC, 1}});
        set(DESTROY_COUNT, 1, 1, 0, true, {{SA, SB, 1}, {SA, SC, 1}});
        set(COUNT, 0, 0, 0, false, {{SA, SA, 2}, {SB, SB, 1}, {SC, SC, 1}});
        set(DOCC, 0, 0, 0, false, {{SA, SA, 1}});
        set(E2D, 0, 0, 0, false, {{SD, SA, 1}});
        set(D2E, 0, 0, 0, false, {{SA, SD, 1}});
        set(FLIP_S0, 2, 1, 1, false, {{SB, SB, r32}, {SC, SC, r32}, {SB, SC, r32}, {SC, SB, r32}});
        ops[FLIP_TO_S2] = ops[FLIP_S0];
        ops[FLIP_TO_S2].action = 2;
        ops[FLIP_TO_S0] = ops[FLIP_S0];
        ops[FLIP_TO_S0].action = -2;
        set(COUNT_FILL, 0, 0, 0, false, {{SA, SA, 2}, {SB, SB, -1}, {SC, SC, -1}, {SB, SC, -1}, {SC, SB, -1}});
        auto pair = [&](int a, int b) { herm[a] = b; herm[b] = a; };
        pair(CREATE_FILL, DESTROY_FILL);
        pair(CREATE, DESTROY);
        pair(E2D, D2E);
        pair(CREATE_FILL_COUNT, DESTROY_FILL_COUNT);
        pair(CREATE_COUNT, DESTROY_COUNT);
        pair(CREATE_FILL_DN, DESTROY_FILL_DN);
        pair(CREATE_UP, DESTROY_UP);
        for (int o : {IDENT, IDENT_FULL, FILL, COUNT, DOCC, COUNT_FILL, FLIP_S0}) {
            selfadj[o] = 1;
            herm[o] = o;
        }
    }
    int conj(int o) const { return herm[o] < 0 ? o : herm[o]; }
    // particle-number change in - out of the operator
    int pdiff(int o) const {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (ops[o].m[i][j] != 0.0) return SQ[i].q[0] - SQ[j].q[0];
        return 0;
    }
    int add(SOp const& s) {
        ops.push_back(s);
        herm.push_back(-1);
        selfadj.push_back(0);
        return int(ops.size()) - 1;
    }
};

// the operator variants that a string position selects from
struct Bundle {
    int up, down, fill_up, fill_down;
};
const Bundle CRE{CREATE_UP, CREATE, CREATE_FILL, CREATE_FILL_DN};
const Bundle DES{DESTROY_UP, DESTROY, DESTROY_FILL, DESTROY_FILL_DN};

struct QTerm {
    int n = 0;
    int pos[4] = {0, 0, 0, 0};
    int op[4] = {0, 0, 0, 0};
    bool operator<(QTerm const& o) const {
        if (n != o.n) return n < o.n;
        for (int i = 0; i < n; ++i) {
            if (pos[i] != o.pos[i]) return pos[i] < o.pos[i];
            if (op[i] != o.op[i]) return op[i] < o.op[i];
        }
        return false;
    }
};

// parity of the permutation (i, j, k, l): number of inversions odd
bool odd_perm(int i, int j, int k, int l) {
    int a[4] = {i, j, k, l}, inv = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = x + 1; y < 4; ++y)
            if (a[x] > a[y]) ++inv;
    return inv & 1;
}

// canonical representative of the 8 index permutations of (ij|kl) (models/chem/util.h:110-117)
std::array<int, 4> align(int i, int j, int k, int l) {
    if (i < j) std::swap(i, j);
    if (k < l) std::swap(k, l);
    if (i < k) { std::swap(i, k); std::swap(j, l); }
    if (i == k && j < l) std::swap(j, l);
    return {i, j, k, l};
}

// The spin-summed Hamiltonian as merged site-ordered operator strings
struct TermSet {
    int L;
    std::map<QTerm, double> strings;           // 2-, 3- and 4-site strings
    std::vector<std::array<double, 16>> site;  // K = 0 single-site terms, m[in*4+out]
    std::vector<char> has_site;
    // gradient layout: the single-site terms kept per operator (site_ops[p][op] = coefficient)
    // instead of summed into one matrix
    bool by_op = false;
    std::vector<std::map<int, double>> site_ops;

    explicit TermSet(int L_, bool by_op_ = false) : L(L_), site(L_), has_site(L_, 0), by_op(by_op_), site_ops(L_) {
        for (auto& s : site) s.fill(0.0);
    }

    void add(int n, const int* pos, const int* op, double c) {
        QTerm t;
        t.n = n;
        int idx[4] = {0, 1, 2, 3};
        std::sort(idx, idx + n, [&](int a, int b) { return pos[a] < pos[b]; });
        for (int x = 0; x < n; ++x) {
            t.pos[x] = pos[idx[x]];
            t.op[x] = op[idx[x]];
        }
        for (int x = 1; x < n; ++x) require(t.pos[x] != t.pos[x - 1], "su2 qc: two operators on one site");
        strings[t] += c;
    }
    void add_site(OpTable const& T, int p, int op, double c) {
        has_site[p] = 1;
        if (by_op) {
            site_ops[p][op] += c;
            return;
        }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) site[p][i * 4 + j] += c * T.ops[op].m[i][j];
    }

    // TermMakerSU2 shapes (term_maker.h:95-225)
    void two_term(double c, int i, int j, int op1, int op2) {
        int pos[2] = {i, j}, op[2] = {op1, op2};
        add(2, pos, op, c);
    }
    void positional(bool sign, double c, int i, int j, int op1, int op1f, int op2, int op2f) {
        int a = i < j ? op1f : op2f, b = i < j ? op2 : op1;
        if (j < i && sign) c = -c;
        int pos[2] = {std::min(i, j), std::max(i, j)}, op[2] = {a, b};
        add(2, pos, op, c);
    }
    void three_term(double c, int pb, int p1, int p2, int bos, int bos_fill, int op1, int op1f, int op2, int op2f) {
        int a = p1 < p2 ? op1f : op2f, b = p1 < p2 ? op2 : op1;
        int bo = ((pb > p1 && pb < p2) || (pb > p2 && pb < p1)) ? bos_fill : bos;
        if (p2 < p1) c = -c;
        int pos[3] = {pb, std::min(p1, p2), std::max(p1, p2)}, op[3] = {bo, a, b};
        add(3, pos, op, c);
    }
    void four_term(int max_two_S, double c, int i, int j, int k, int l, Bundle const& bi, Bundle const& bk) {
        if (odd_perm(i, j, k, l)) c = -c;
        std::pair<int, Bundle const*> s[4] = {{i, &bi}, {j, &bi}, {k, &bk}, {l, &bk}};
        std::sort(s, s + 4, [](auto const& x, auto const& y) { return x.first < y.first; });
        int pos[4], op[4];
        for (int x = 0; x < 4; ++x) pos[x] = s[x].first;
        if (max_two_S == 2) {
            op[0] = s[0].second->fill_up;
            op[1] = s[1].second->up;
            op[2] = s[2].second->fill_down;
            op[3] = s[3].second->down;
        } else {
            op[0] = s[0].second->fill_up;
            op[1] = s[1].second->down;
            op[2] = s[2].second->fill_up;
            op[3] = s[3].second->down;
        }
        add(4, pos, op, c);
    }

    // SpinSumSU2 (term_maker.h:228-460): sum_st c+_i,s c+_k,t c_l,t c_j,s for an index pattern
    void v_four(double V, int i, int k, int l, int j) {
        auto a = align(i, j, k, l);
        int j_ = a[1], k_ = a[2], l_ = a[3];
        const double r3 = std::sqrt(3.0);
        if (k_ > l_ && l_ > j_) {
            four_term(2, -r3 * V, i, k, l, j, CRE, DES);
            four_term(1, V, i, k, l, j, CRE, DES);
        } else if (k_ > j_ && j_ > l_) {
            double loc = odd_perm(i, k, l, j) ? -V : V;
            four_term(2, r3 * loc, i, k, l, j, CRE, DES);
            four_term(1, loc, i, k, l, j, CRE, DES);
        } else if (j_ > k_ && k_ > l_) {
            four_term(1, 2.0 * V, i, k, l, j, CRE, DES);
        } else {
            throw Error(MQ_ERR_INVALID, "su2 qc: unexpected index arrangement of a 4-index term");
        }
    }
    void v_two(double V, int i, int k, int l, int j) {
        const double r2 = std::sqrt(2.0), r3 = std::sqrt(3.0);
        if (i == j && k == l && j != k) {
            two_term(V, i, k, COUNT, COUNT);
        } else if (i == k && j == l && j != k) {
            two_term(2.0 * V, i, j, E2D, D2E);
        } else if (i == l && j == k && i != j) {
            positional(false, r3 * V, i, j, FLIP_TO_S0, FLIP_TO_S2, FLIP_TO_S0, FLIP_TO_S2);
            two_term(-0.5 * V, i, j, COUNT, COUNT);
        } else if ((i == k && k == l) || (k == l && l == j) || (i == l && l == j) || (i == k && k == j)) {
            int s, p;
            if (i == k && k == l) { s = i; p = j; }
            else if (k == l && l == j) { s = j; p = i; }
            else if (i == l && l == j) { s = i; p = k; }
            else { s = i; p = l; }
            if (i == k)  // one lone annihilator
                positional(true, r2 * V, s, p, CREATE_COUNT, CREATE_FILL_COUNT, DESTROY, DESTROY_FILL);
            else         // one lone creator
                positional(true, -r2 * V, s, p, DESTROY_COUNT, DESTROY_FILL_COUNT, CREATE, CREATE_FILL);
        } else {
            throw Error(MQ_ERR_INVALID, "su2 qc: unexpected index arrangement of a 2-site term");
        }
    }
    void v_three(double V, int i, int k, int l, int j) {
        const double r2 = std::sqrt(2.0), r3 = std::sqrt(3.0);
        if (i == j || k == l) {
            int same = i;
            if (k == l) { same = k; k = i; l = j; }
            three_term(r2 * V, same, k, l, COUNT, COUNT_FILL, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
            return;
        }
        if (i == k) {
            three_term(-r2 * V, i, std::min(l, j), std::max(l, j), E2D, E2D, DESTROY, DESTROY_FILL, DESTROY, DESTROY_FILL);
        }
        if (j == l) {
            three_term(-r2 * V, j, std::min(i, k), std::max(i, k), D2E, D2E, CREATE, CREATE_FILL, CREATE, CREATE_FILL);
        }
        if (j == k || i == l) {
            int same, p1, p2;
            if (j == k) { same = j; p1 = l; p2 = i; }
            else { same = i; p1 = j; p2 = k; }
            double phase = odd_perm(i, k, l, j) ? -1.0 : 1.0;
            if (same < std::min(p1, p2)) {
                three_term(phase * r3 * V, same, p1, p2, FLIP_TO_S2, FLIP_TO_S2, CREATE, CREATE_FILL_DN, DESTROY,
                           DESTROY_FILL_DN);
                three_term(-0.5 * r2 * V, same, p1, p2, COUNT, COUNT, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
            } else if (same > std::max(p1, p2)) {
                three_term(phase * r3 * V, same, p1, p2, FLIP_TO_S0, FLIP_TO_S0, CREATE_UP, CREATE_FILL, DESTROY_UP,
                           DESTROY_FILL);
                three_term(-0.5 * r2 * V, same, p1, p2, COUNT, COUNT, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
            } else {
                three_term(phase * r3 * V, same, p1, p2, FLIP_S0, FLIP_S0, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
                three_term(-0.5 * r2 * V, same, p1, p2, COUNT_FILL, COUNT_FILL, CREATE, CREATE_FILL, DESTROY,
                           DESTROY_FILL);
            }
        }
    }

    // one canonical integral (i >= j, k >= l, (i,j) >= (k,l)); one-electron terms have k = l = -1
    // (create_terms, su2u1/model.hpp:347-588)
    void integral(OpTable const& T, int i, int j, int k, int l, double V) {
        const double r2 = std::sqrt(2.0);
        if (k < 0) {
            if (i == j) {
                add_site(T, i, COUNT, V);
            } else {
                // sum_s h_ij (c+_is c_js + c+_js c_is) as two spin-coupled creator/annihilator pairs
                positional(true, r2 * V, j, i, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
                positional(true, r2 * V, i, j, CREATE, CREATE_FILL, DESTROY, DESTROY_FILL);
            }
            return;
        }
        if (i == j && j == k && k == l) {
            add_site(T, i, DOCC, V);
        } else if ((i == j && j == k && k != l) || (i != j && j == k && k == l)) {
            int s = i == j ? i : l, p = i == j ? l : i;
            v_two(V, s, s, s, p);
            v_two(V, s, p, s, s);
        } else if (i == j && k == l && j != k) {
            v_two(V, i, k, k, i);
        } else if (i == k && j == l && i != j) {
            v_two(0.5 * V, i, i, j, j);
            v_two(0.5 * V, j, j, i, i);
            v_two(V, i, j, i, j);
        } else if ((i == j && j != k && k != l) || (k == l && i != j && j != k)) {
            if (i == j) {
                v_three(V, i, k, l, i);
                v_three(V, i, l, k, i);
            } else {
                v_three(V, i, k, k, j);
                v_three(V, j, k, k, i);
            }
        } else if (((i == k && j != l) || j == k || (j == l && i != k)) && (i != j && k != l)) {
            v_three(V, i, k, l, j);
            v_three(V, i, l, k, j);
            v_three(V, j, k, l, i);
            v_three(V, j, l, k, i);
        } else if (i != j && j != k && k != l && i != k && j != l) {
            v_four(V, i, k, l, j);
            v_four(V, i, l, k, j);
            v_four(V, j, k, l, i);
            v_four(V, j, l, k, i);
        } else {
            throw Error(MQ_ERR_INVALID, "su2 qc: integral index pattern not covered");
        }
    }
};

// ---- MPO assembly ----------------------------------------------------------------------------

// bond-state key: kind 0 = unfinished (trivial left), 1 = applied operators, 2 = finished
// (trivial right), 3 = pending operators
struct Key {
    uint8_t kind = 0, n = 0;
    int16_t pos[2] = {0, 0};
    int16_t op[2] = {0, 0};
    bool operator<(Key const& o) const {
        if (kind != o.kind) return kind < o.kind;
        if (n != o.n) return n < o.n;
        for (int i = 0; i < n; ++i) {
            if (pos[i] != o.pos[i]) return pos[i] < o.pos[i];
            if (op[i] != o.op[i]) return op[i] < o.op[i];
        }
        return false;
    }
    bool operator==(Key const& o) const { return !(*this < o) && !(o < *this); }
};
Key key_of(uint8_t kind, int n, const int* pos, const int* op) {
    Key k;
    k.kind = kind;
    k.n = uint8_t(n);
    for (int i = 0; i < n; ++i) { k.pos[i] = int16_t(pos[i]); k.op[i] = int16_t(op[i]); }
    return k;
}
const Key TL{0, 0, {0, 0}, {0, 0}};
const Key TR{2, 0, {0, 0}, {0, 0}};

struct Cell {
    std::vector<std::pair<int, double>> terms;
    void add(int op, double s) {
        for (auto& x : terms)
            if (x.first == op) { x.second += s; return; }
        terms.push_back({op, s});
    }
};

// per-product-state helpers of the two-site index
struct Phys2 {
    std::vector<Charge> charges;  // descending
    std::vector<uint32_t> sizes;
    int sector[16], offset[16];
    Phys2() {
        std::map<Charge, uint32_t, std::greater<Charge>> sz;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) sz[SQ[a] + SQ[b]] += 1;
        for (auto& kv : sz) { charges.push_back(kv.first); sizes.push_back(kv.second); }
        std::map<Charge, int> fill;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                Charge c = SQ[a] + SQ[b];
                sector[a * 4 + b] = int(std::find(charges.begin(), charges.end(), c) - charges.begin());
                offset[a * 4 + b] = fill[c]++;
            }
    }
};

}  // namespace

struct SU2QC {
    int L = 0;
    OpTable T;
    std::vector<std::map<std::pair<Key, Key>, Cell>> prempo;
    std::vector<FlatMPO> site_mpo;
    std::vector<std::vector<int>> site_ops;  // per site: FlatMPO op index -> table op
    std::vector<std::unique_ptr<FlatMPO>> two;
    std::vector<uint32_t> bond_dims;
    // gradient layout (build with grad = true): every canonical integral's terms are kept, zero
    // coefficients included, so the chain's structure does not depend on the integral values;
    // single-site terms are one MPO term per operator; and each string remembers the MPO term its
    // coefficient sits on (the term scale is the sum of the coefficients of the strings on it)
    bool grad = false;
    std::map<QTerm, std::tuple<int, Key, Key, int>> detach_at;  // string -> (site, keys, op)
    std::map<QTerm, std::pair<int, uint32_t>> string_term;      // string -> (site, term)
    std::vector<std::map<int, uint32_t>> site_term;             // site -> op -> term

    // ---- key properties ----
    int key_spin(Key const& k) const {
        int s = 0;
        for (int i = 0; i < k.n; ++i) s += T.ops[k.op[i]].action;
        return k.kind == 3 ? -s : s;
    }
    Key conj_key(Key k) const {
        for (int i = 0; i < k.n; ++i) k.op[i] = int16_t(T.conj(k.op[i]));
        return k;
    }
    // a key whose operators are all self-adjoint, or a single rank-2 spin-flip operator (the spin
    // operator of one site, its own adjoint up to the reduced-element correction)
    bool self_adjoint(Key const& k) const {
        if (k.n == 0) return false;
        if (k.n == 1 && (k.op[0] == FLIP_TO_S2 || k.op[0] == FLIP_TO_S0)) return true;
        for (int i = 0; i < k.n; ++i)
            if (!T.selfadj[k.op[i]]) return false;
        return true;
    }
    // phase of the Hermitian partner relation of a key: (-1) per creator of rank 1/2, times
    // (-1)^S for a coupled operator pair of spin S.  For single operators and applied pairs this is
    // the reference's rule (tagged_mpo_maker_optim.hpp:620-690, which assigns the pair's phases
    // from the lower-numbered key); pending pairs, which the reference's keying never forms, follow
    // the same rule (checked block by block against uncompressed boundaries)
    int key_phase(Key const& k) const {
        int ph = 1;
        for (int i = 0; i < k.n; ++i)
            if (T.ops[k.op[i]].K == 1 && T.pdiff(k.op[i]) == -1) ph = -ph;
        if (k.n == 2 && ((key_spin(k) / 2) & 1)) ph = -ph;
        return ph;
    }

    // ---- prempo insertion ----
    int fill_for(Key const& k, int pending_ferm) const {
        if (pending_ferm & 1) return FILL;
        return key_spin(k) > 1 ? IDENT_FULL : IDENT;
    }
    void insert_fill(int from, int to, Key const& k, int op) {
        for (int p = from; p < to; ++p) {
            auto& cell = prempo[p][{k, k}];
            if (cell.terms.empty()) cell.terms.push_back({op, 1.0});
            else require(cell.terms.size() == 1 && cell.terms[0].first == op, "su2 qc: conflicting fill operator");
        }
    }
    void attach(int p, Key const& a, Key const& b, int op) {
        auto& cell = prempo[p][{a, b}];
        if (cell.terms.empty()) cell.terms.push_back({op, 1.0});
    }
    void detach(QTerm const& t, int p, Key const& a, Key const& b, int op, double c) {
        prempo[p][{a, b}].add(op, c);
        if (grad) detach_at[t] = std::make_tuple(p, a, b, op);
    }

    void add_string(QTerm const& t, double c) {
        auto nferm = [&](int from) {
            int f = 0;
            for (int x = from; x < t.n; ++x) f += T.ops[t.op[x]].fermionic;
            return f;
        };
        const int* P = t.pos;
        const int* O = t.op;
        if (t.n == 2) {
            Key k = key_of(3, 1, P + 1, O + 1);
            detach(t, P[0], TL, k, O[0], c);
            insert_fill(P[0] + 1, P[1], k, fill_for(k, nferm(1)));
            attach(P[1], k, TR, O[1]);
        } else if (t.n == 3) {
            Key k1 = key_of(1, 1, P, O);
            attach(P[0], TL, k1, O[0]);
            insert_fill(P[0] + 1, P[1], k1, fill_for(k1, nferm(1)));
            Key k2 = key_of(3, 1, P + 2, O + 2);
            detach(t, P[1], k1, k2, O[1], c);
            insert_fill(P[1] + 1, P[2], k2, fill_for(k2, nferm(2)));
            attach(P[2], k2, TR, O[2]);
        } else if (P[1] + P[2] < L - 1) {
            Key k1 = key_of(1, 1, P, O);
            attach(P[0], TL, k1, O[0]);
            insert_fill(P[0] + 1, P[1], k1, fill_for(k1, nferm(1)));
            Key k2 = key_of(1, 2, P, O);
            attach(P[1], k1, k2, O[1]);
            insert_fill(P[1] + 1, P[2], k2, fill_for(k2, nferm(2)));
            Key k3 = key_of(3, 1, P + 3, O + 3);
            detach(t, P[2], k2, k3, O[2], c);
            insert_fill(P[2] + 1, P[3], k3, fill_for(k3, nferm(3)));
            attach(P[3], k3, TR, O[3]);
        } else {
            Key k1 = key_of(1, 1, P, O);
            attach(P[0], TL, k1, O[0]);
            insert_fill(P[0] + 1, P[1], k1, fill_for(k1, nferm(1)));
            Key k2 = key_of(3, 2, P + 2, O + 2);
            detach(t, P[1], k1, k2, O[1], c);
            insert_fill(P[1] + 1, P[2], k2, fill_for(k2, nferm(2)));
            Key k3 = key_of(3, 1, P + 3, O + 3);
            attach(P[2], k2, k3, O[2]);
            insert_fill(P[2] + 1, P[3], k3, fill_for(k3, nferm(3)));
            attach(P[3], k3, TR, O[3]);
        }
    }

    // the canonical integrals (i >= j, k >= l, (i, j) >= (k, l); one-electron ones with k = l = -1)
    // in the builder's order
    template <class F>
    void for_each_integral(F f) const {
        for (int i = 0; i < L; ++i)
            for (int j = 0; j <= i; ++j) f(i, j, -1, -1);
        for (int i = 0; i < L; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = 0; k <= i; ++k)
                    for (int l = 0; l <= k; ++l) {
                        if (k == i && l > j) continue;
                        f(i, j, k, l);
                    }
    }
    static size_t integral_pos(int L, int i, int j, int k, int l) {
        return k < 0 ? size_t(i) * L + j : ((size_t(i) * L + j) * L + k) * L + l;
    }

    void build(const double* h1, const double* eri, double cutoff) {
        TermSet S(L, grad);
        for_each_integral([&](int i, int j, int k, int l) {
            double v = k < 0 ? h1[integral_pos(L, i, j, k, l)] : eri[integral_pos(L, i, j, k, l)];
            if (grad || std::fabs(v) > cutoff) S.integral(T, i, j, k, l, v);
        });
        prempo.assign(L, {});
        int rightmost_left = 0, leftmost_right = L - 1;
        for (auto const& kv : S.strings) {
            if (kv.second == 0.0 && !grad) continue;
            QTerm const& t = kv.first;
            rightmost_left = std::max(rightmost_left, t.pos[0]);
            leftmost_right = std::min(leftmost_right, t.pos[t.n - 1]);
            add_string(t, kv.second);
        }
        for (int p = 0; p < L; ++p)
            if (S.has_site[p] && grad) {
                rightmost_left = std::max(rightmost_left, p);
                leftmost_right = std::min(leftmost_right, p);
                for (auto const& kv : S.site_ops[p]) prempo[p][{TL, TR}].add(kv.first, kv.second);
            } else if (S.has_site[p]) {
                rightmost_left = std::max(rightmost_left, p);
                leftmost_right = std::min(leftmost_right, p);
                SOp s;
                bool any = false;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) {
                        s.m[i][j] = S.site[p][i * 4 + j];
                        any |= s.m[i][j] != 0.0;
                    }
                if (any) prempo[p][{TL, TR}].add(T.add(s), 1.0);
            }
        for (int p = 0; p < rightmost_left; ++p) attach(p, TL, TL, IDENT);
        insert_fill(leftmost_right + 1, L, TR, IDENT);
        assemble();
    }

    // one table op as SU2U1 site-operator blocks of a FlatMPO (1x1 blocks per sector pair)
    uint32_t append_op(FlatMPO& F, int o) {
        SOp const& s = T.ops[o];
        uint32_t idx = uint32_t(F.op_spin.size());
        F.op_spin.push_back(s.K);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                if (s.m[i][j] == 0.0) continue;
                F.block_charges.insert(F.block_charges.end(), {SQ[i].q[0], SQ[i].q[1], SQ[j].q[0], SQ[j].q[1]});
                F.entry_rc.push_back(0);
                F.entry_rc.push_back(0);
                F.entry_spins.push_back(abs_spin(i));
                F.entry_spins.push_back(abs_spin(j));
                F.entry_coeff.push_back(s.m[i][j]);
                F.block_entry_ptr.push_back(int64_t(F.entry_coeff.size()));
            }
        F.op_block_ptr.push_back(int64_t(F.block_charges.size() / 4));
        return idx;
    }

    void assemble() {
        std::map<std::tuple<int, Key, Key, int>, uint32_t> flat_of;  // grad: (site, keys, op) -> term
        site_mpo.assign(L, FlatMPO());
        site_ops.assign(L, {});
        bond_dims.assign(L + 1, 1);
        std::map<Key, uint32_t> left;
        left[TL] = 0;
        std::vector<int32_t> lspin{0}, lphase{1};
        std::vector<uint32_t> lherm{MQ_NO_PARTNER};
        for (int p = 0; p < L; ++p) {
            std::vector<Key> k2s;
            for (auto const& kv : prempo[p]) k2s.push_back(kv.first.second);
            std::sort(k2s.begin(), k2s.end());
            k2s.erase(std::unique(k2s.begin(), k2s.end()), k2s.end());
            std::map<Key, uint32_t> right;
            uint32_t r = 2;
            for (auto const& k : k2s) {
                if (p == L - 1) right[k] = 0;
                else if (k.kind == 0) right[k] = 0;
                else if (k.kind == 2) right[k] = 1;
                else right[k] = r++;
            }
            uint32_t rdim = (p == L - 1) ? 1 : std::max<uint32_t>(r, 2);
            FlatMPO& F = site_mpo[p];
            F.row_dim = bond_dims[p];
            F.col_dim = rdim;
            std::map<std::pair<uint32_t, uint32_t>, std::pair<Key, Key> const*> cells;
            for (auto const& kv : prempo[p]) {
                auto li = left.find(kv.first.first);
                require(li != left.end(), "su2 qc: left key not found");
                cells[{li->second, right[kv.first.second]}] = &kv.first;
            }
            std::map<int, uint32_t> op_index;
            for (auto const& c : cells)
                for (auto const& t : prempo[p].at(*c.second).terms) {
                    if (t.second == 0.0 && !grad) continue;
                    auto it = op_index.find(t.first);
                    if (it == op_index.end()) {
                        it = op_index.emplace(t.first, append_op(F, t.first)).first;
                        site_ops[p].push_back(t.first);
                    }
                    if (grad) flat_of[std::make_tuple(p, c.second->first, c.second->second, t.first)] =
                        uint32_t(F.term_op.size());
                    F.term_b.push_back(c.first.first);
                    F.term_b.push_back(c.first.second);
                    F.term_op.push_back(it->second);
                    F.term_scale.push_back(t.second);
                }
            F.lspin = lspin;
            F.lherm = lherm;
            F.lphase = lphase;
            std::vector<int32_t> rspin(rdim, 0), rphase(rdim, 1);
            std::vector<uint32_t> rherm(rdim, MQ_NO_PARTNER);
            if (p < L - 1) {
                for (auto const& kv : right) {
                    rspin[kv.second] = key_spin(kv.first);
                    if (kv.first.kind == 0 || kv.first.kind == 2) continue;
                    if (self_adjoint(kv.first)) continue;
                    Key ck = conj_key(kv.first);
                    if (ck == kv.first) continue;
                    auto it = right.find(ck);
                    if (it == right.end()) continue;
                    uint32_t romeo = kv.second, julia = it->second;
                    if (romeo >= julia) continue;
                    rherm[romeo] = julia;
                    rherm[julia] = romeo;
                    rphase[romeo] = key_phase(kv.first);
                    rphase[julia] = key_phase(it->first);
                }
                for (auto const& kv : right)
                    if (kv.first.kind != 0 && kv.first.kind != 2 && self_adjoint(kv.first)) rherm[kv.second] = kv.second;
            }
            F.rspin = rspin;
            F.rherm = rherm;
            F.rphase = rphase;
            bond_dims[p + 1] = rdim;
            left = right;
            lspin = rspin;
            lherm = rherm;
            lphase = rphase;
        }
        prempo.clear();
        two.resize(L > 0 ? L - 1 : 0);
        if (!grad) return;
        for (auto const& kv : detach_at) {
            auto it = flat_of.find(kv.second);
            require(it != flat_of.end(), "su2 qc: a string's coefficient term is missing");
            string_term[kv.first] = {std::get<0>(kv.second), it->second};
        }
        detach_at.clear();
        site_term.assign(L, {});
        for (auto const& kv : flat_of)
            if (std::get<1>(kv.first) == TL && std::get<2>(kv.first) == TR)
                site_term[std::get<0>(kv.first)][std::get<3>(kv.first)] = kv.second;
    }

    // dE/d(integral) of every canonical integral from dE/d(term scale) of every site MPO term:
    // the builder is linear in the integrals, so the derivative of a term scale with respect to one
    // integral is what that integral alone (at 1) adds to it
    void integral_gradient(std::vector<const double*> const& e, double* dh, double* deri) const {
        require(grad, "su2 qc: integral gradients need the gradient layout");
        require(e.size() == size_t(L), "su2 qc: one term-gradient array per site expected");
        std::fill(dh, dh + size_t(L) * L, 0.0);
        std::fill(deri, deri + size_t(L) * L * L * L, 0.0);
        for_each_integral([&](int i, int j, int k, int l) {
            TermSet S(L, true);
            S.integral(T, i, j, k, l, 1.0);
            double g = 0.0;
            for (auto const& kv : S.strings) {
                auto it = string_term.find(kv.first);
                require(it != string_term.end(), "su2 qc: an integral's string is not in the chain");
                g += kv.second * e[size_t(it->second.first)][it->second.second];
            }
            for (int p = 0; p < L; ++p)
                for (auto const& kv : S.site_ops[p]) {
                    auto it = site_term[size_t(p)].find(kv.first);
                    require(it != site_term[size_t(p)].end(), "su2 qc: a single-site term is not in the chain");
                    g += kv.second * e[size_t(p)][it->second];
                }
            (k < 0 ? dh : deri)[integral_pos(L, i, j, k, l)] = g;
        });
    }

    // ---- two-site MPO ----
    // SU(2) Kronecker product of two site operators between bond spins j (left), jm (middle),
    // jp (right) at total rank k (site_operator_algorithms.h:214-335); m[in][out] over product
    // states a*4+b
    // mod_coupling and the 6j factor of kron take spins of at most 4 (as 2S): memoised
    mutable std::unordered_map<uint64_t, double> cpl_cache;
    double cached(int kind, int a, int b, int c, int d, int e, int f, int g = 0, int h = 0, int i = 0) const {
        uint64_t key = uint64_t(kind);
        for (int v : {a, b, c, d, e, f, g, h, i}) key = key * 8 + uint64_t(v);
        auto it = cpl_cache.find(key);
        if (it != cpl_cache.end()) return it->second;
        double v = kind == 0 ? mod_coupling(a, b, c, d, e, f, g, h, i) : wigner_6j(a, b, c, d, e, f);
        cpl_cache.emplace(key, v);
        return v;
    }
    void kron(int oa, int ob, int j, int jm, int jp, int k, std::array<double, 256>& out) const {
        SOp A = T.ops[oa], B = T.ops[ob];
        // a rank-0 partner of a rank > 0 operator acts as the full identity on the singly
        // occupied pair
        if (A.K > 0 && B.K == 0 && B.m[SB][SC] == 0.0) { B.m[SB][SC] = 1.0; B.m[SC][SB] = 1.0; }
        if (A.K == 0 && B.K > 0 && A.m[SB][SC] == 0.0) { A.m[SB][SC] = 1.0; A.m[SC][SB] = 1.0; }
        int k1 = A.K, k2 = B.K;
        double f = std::sqrt((jm + 1.0) * (k + 1.0)) * cached(1, j, jp, k, k2, k1, jm);
        if (((j + jp + k1 + k2) / 2) % 2) f = -f;
        out.fill(0.0);
        if (f == 0.0) return;
        for (int ia = 0; ia < 4; ++ia)
            for (int oa2 = 0; oa2 < 4; ++oa2) {
                double va = A.m[ia][oa2];
                if (va == 0.0) continue;
                for (int ib = 0; ib < 4; ++ib)
                    for (int ob2 = 0; ob2 < 4; ++ob2) {
                        double vb = B.m[ib][ob2];
                        if (vb == 0.0) continue;
                        double c = cached(0, abs_spin(ia), abs_spin(ib), product_spin(ia, ib), k1, k2, k,
                                          abs_spin(oa2), abs_spin(ob2), product_spin(oa2, ob2));
                        out[(ia * 4 + ib) * 16 + (oa2 * 4 + ob2)] += va * vb * c * f;
                    }
            }
    }

    FlatMPO make_twosite(int p) {
        FlatMPO const& A = site_mpo[p];
        FlatMPO const& B = site_mpo[p + 1];
        Phys2 P2;
        FlatMPO Tw;
        Tw.row_dim = A.row_dim;
        Tw.col_dim = B.col_dim;
        std::vector<std::vector<size_t>> brow(B.row_dim);
        for (size_t t = 0; t < B.term_op.size(); ++t) brow[B.term_b[2 * t]].push_back(t);
        // (b0, b2) -> k -> list of (op of A, op of B, scale)
        std::map<std::pair<uint32_t, uint32_t>, std::map<int, std::vector<std::tuple<int, int, double>>>> cells;
        for (size_t ta = 0; ta < A.term_op.size(); ++ta) {
            uint32_t b0 = A.term_b[2 * ta], b1 = A.term_b[2 * ta + 1];
            int oa = site_ops[p][A.term_op[ta]];
            for (size_t tb : brow[b1]) {
                uint32_t b2 = B.term_b[2 * tb + 1];
                int ob = site_ops[p + 1][B.term_op[tb]];
                int j = A.lspin[b0], jp = B.rspin[b2];
                int k1 = T.ops[oa].K, k2 = T.ops[ob].K;
                for (int k = std::abs(k1 - k2); k <= k1 + k2; k += 2) {
                    if (!(jp >= std::abs(k - j) && jp <= std::abs(k + j))) continue;
                    cells[{b0, b2}][k].push_back(std::make_tuple(ta, int(tb), A.term_scale[ta] * B.term_scale[tb]));
                }
            }
        }
        // two-site operators deduplicated by rank and entries
        std::map<std::pair<int, std::vector<std::pair<int, double>>>, uint32_t> opidx;
        auto register_op = [&](int k, std::array<double, 256> const& m) -> int64_t {
            std::vector<std::pair<int, double>> ents;
            for (int x = 0; x < 256; ++x)
                if (m[x] != 0.0) ents.push_back({x, m[x]});
            if (ents.empty()) return -1;
            auto key = std::make_pair(k, ents);
            auto it = opidx.find(key);
            if (it != opidx.end()) return it->second;
            uint32_t idx = uint32_t(Tw.op_spin.size());
            Tw.op_spin.push_back(k);
            // blocks by (in sector, out sector), entries in (in offset, out offset) order
            std::map<std::pair<int, int>, std::vector<std::tuple<int, int, int, int, double>>> blocks;
            for (auto const& e : ents) {
                int in = e.first / 16, out = e.first % 16;
                blocks[{P2.sector[in], P2.sector[out]}].push_back(std::make_tuple(
                    P2.offset[in], P2.offset[out], product_spin(in / 4, in % 4), product_spin(out / 4, out % 4), e.second));
            }
            for (auto& bl : blocks) {
                std::sort(bl.second.begin(), bl.second.end());
                Charge lc = P2.charges[bl.first.first], rc = P2.charges[bl.first.second];
                Tw.block_charges.insert(Tw.block_charges.end(), {lc.q[0], lc.q[1], rc.q[0], rc.q[1]});
                for (auto const& e : bl.second) {
                    Tw.entry_rc.push_back(uint32_t(std::get<0>(e)));
                    Tw.entry_rc.push_back(uint32_t(std::get<1>(e)));
                    Tw.entry_spins.push_back(std::get<2>(e));
                    Tw.entry_spins.push_back(std::get<3>(e));
                    Tw.entry_coeff.push_back(std::get<4>(e));
                }
                Tw.block_entry_ptr.push_back(int64_t(Tw.entry_coeff.size()));
            }
            Tw.op_block_ptr.push_back(int64_t(Tw.block_charges.size() / 4));
            opidx.emplace(key, idx);
            return idx;
        };
        std::array<double, 256> m, acc;
        for (auto const& c : cells) {
            uint32_t b0 = c.first.first, b2 = c.first.second;
            int j = A.lspin[b0], jp = B.rspin[b2];
            for (auto const& kl : c.second) {
                int k = kl.first;
                auto const& lst = kl.second;
                if (lst.size() == 1) {
                    int ta = std::get<0>(lst[0]), tb = std::get<1>(lst[0]);
                    kron(site_ops[p][A.term_op[ta]], site_ops[p + 1][B.term_op[tb]], j, A.rspin[A.term_b[2 * ta + 1]],
                         jp, k, m);
                    int64_t o = register_op(k, m);
                    if (o < 0) continue;
                    Tw.term_b.push_back(b0);
                    Tw.term_b.push_back(b2);
                    Tw.term_op.push_back(uint32_t(o));
                    Tw.term_scale.push_back(std::get<2>(lst[0]));
                } else {
                    acc.fill(0.0);
                    for (auto const& e : lst) {
                        int ta = std::get<0>(e), tb = std::get<1>(e);
                        kron(site_ops[p][A.term_op[ta]], site_ops[p + 1][B.term_op[tb]], j,
                             A.rspin[A.term_b[2 * ta + 1]], jp, k, m);
                        for (int x = 0; x < 256; ++x) acc[x] += std::get<2>(e) * m[x];
                    }
                    int64_t o = register_op(k, acc);
                    if (o < 0) continue;
                    Tw.term_b.push_back(b0);
                    Tw.term_b.push_back(b2);
                    Tw.term_op.push_back(uint32_t(o));
                    Tw.term_scale.push_back(1.0);
                }
            }
        }
        Tw.lspin = A.lspin;
        Tw.rspin = B.rspin;
        Tw.lherm = A.lherm;
        Tw.lphase = A.lphase;
        Tw.rherm = B.rherm;
        Tw.rphase = B.rphase;
        return Tw;
    }
};

void SU2QCDeleter::operator()(SU2QC* p) const { delete p; }

std::unique_ptr<SU2QC, SU2QCDeleter> su2qc_build(int L, const double* h1, const double* eri, double cutoff,
                                                 bool gradient_layout) {
    std::unique_ptr<SU2QC, SU2QCDeleter> q(new SU2QC());
    q->L = L;
    q->grad = gradient_layout;
    q->build(h1, eri, cutoff);
    return q;
}
void su2qc_integral_gradient(SU2QC const& q, std::vector<const double*> const& term_grads, double* dh, double* deri) {
    q.integral_gradient(term_grads, dh, deri);
}
int su2qc_length(SU2QC const& q) { return q.L; }
std::vector<uint32_t> const& su2qc_bond_dims(SU2QC const& q) { return q.bond_dims; }
FlatMPO const& su2qc_site(SU2QC& q, int p) { return q.site_mpo[p]; }
FlatMPO const& su2qc_twosite(SU2QC& q, int p) {
    if (!q.two[p]) q.two[p].reset(new FlatMPO(q.make_twosite(p)));
    return *q.two[p];
}
void su2qc_phys(int twosite, std::vector<int32_t>& charges, std::vector<uint32_t>& sizes) {
    charges.clear();
    sizes.clear();
    if (!twosite) {
        for (int s = 0; s < 4; ++s) {
            charges.push_back(SQ[s].q[0]);
            charges.push_back(SQ[s].q[1]);
            sizes.push_back(1);
        }
        return;
    }
    Phys2 P;
    for (size_t i = 0; i < P.charges.size(); ++i) {
        charges.push_back(P.charges[i].q[0]);
        charges.push_back(P.charges[i].q[1]);
        sizes.push_back(P.sizes[i]);
    }
}

}  // namespace mq
