// Quantum-chemistry Hamiltonian MPO (TwoU1), the input producer of the contraction path.
//
//   H = E_core + sum_pq h_pq sum_s c+_ps c_qs + 1/2 sum_pqrs (pq|rs) sum_st c+_ps c+_rt c_st c_qs
//
// Construction (reference: models/chem/2u1/model.hpp create_terms, generate_mpo/
// tagged_mpo_maker_optim.hpp add_2term/add_3term/add_4term/finalize/create_mpo):
//   * every fermionic operator string is written in the Jordan-Wigner tensor-product form: the
//     operator on site q is the ordered product over the string of F (q before the operator's
//     site), the local operator (q at its site) or 1 (q after it); F = (-1)^n.  Products of such
//     strings factorise site by site, so no sign bookkeeping is needed beyond that.
//   * strings with the same (site, local operator) content are merged;
//   * MPO bond states are keyed like the reference's prempo keys: after the first operator of a
//     3- or 4-site string the key is the applied operator, after the second operator of a 4-site
//     string the applied pair, and the coefficient sits on the second-to-last operator whose key
//     is the pending last operator ("detach").  Fill operators (F or 1 by the parity of the pending
//     operators) run between operator sites; identity paths carry the unfinished (index 0) and
//     finished (index 1) states.  Bond dimension O(L^2).
//   * Hermitian partners: the key of the transposed operators is the partner (phase 1 for abelian
//     groups, tagged_mpo_maker_optim.hpp:597-606); keys of self-adjoint operators are self-adjoint.
// Site basis (descending TwoU1 order): 0 = (1,1) up+down, 1 = (1,0) up, 2 = (0,1) down, 3 = (0,0).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "core.hpp"
#include "qc_mpo.hpp"
#include "qcmaquis_amd.h"

namespace mq {
namespace {

typedef std::array<double, 16> Mat4;  // M[out*4 + in]

Mat4 mul(Mat4 const& a, Mat4 const& b) {
    Mat4 c{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            c[i * 4 + j] = s;
        }
    return c;
}
Mat4 transpose(Mat4 const& a) {
    Mat4 t{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t[i * 4 + j] = a[j * 4 + i];
    return t;
}
bool is_zero(Mat4 const& a) {
    for (double v : a)
        if (v != 0.0) return false;
    return true;
}

const Charge SITE_Q[4] = {Charge(1, 1), Charge(1, 0), Charge(0, 1), Charge(0, 0)};

Mat4 op_ident() { Mat4 m{}; for (int i = 0; i < 4; ++i) m[i * 5] = 1; return m; }
Mat4 op_fill() { Mat4 m{}; m[0] = 1; m[5] = -1; m[10] = -1; m[15] = 1; return m; }
// c+_up: |0> -> |up>, |dn> -> |updn>;  c+_dn: |0> -> |dn>, |up> -> -|updn>
Mat4 op_cdag(int spin) {
    Mat4 m{};
    if (spin == 0) { m[1 * 4 + 3] = 1; m[0 * 4 + 2] = 1; }
    else { m[2 * 4 + 3] = 1; m[0 * 4 + 1] = -1; }
    return m;
}
Mat4 op_c(int spin) { return transpose(op_cdag(spin)); }

struct Hasher {
    size_t operator()(Mat4 const& m) const {
        size_t h = 1469598103934665603ull;
        for (double v : m) {
            uint64_t b;
            std::memcpy(&b, &v, 8);
            h = (h ^ b) * 1099511628211ull;
        }
        return h;
    }
};

// ---- tagged operator table ----
struct OpTable {
    std::vector<Mat4> ops;
    std::unordered_map<Mat4, uint32_t, Hasher> index;
    uint32_t tag(Mat4 const& m) {
        auto it = index.find(m);
        if (it != index.end()) return it->second;
        uint32_t t = uint32_t(ops.size());
        ops.push_back(m);
        index[m] = t;
        return t;
    }
    bool fermionic(uint32_t t) const {  // odd particle-number change
        Mat4 const& m = ops[t];
        for (int o = 0; o < 4; ++o)
            for (int i = 0; i < 4; ++i)
                if (m[o * 4 + i] != 0.0) {
                    int dn = (SITE_Q[o].q[0] + SITE_Q[o].q[1]) - (SITE_Q[i].q[0] + SITE_Q[i].q[1]);
                    return (dn & 1) != 0;
                }
        return false;
    }
};

// bond-state key: kind 0 = unfinished (trivial left), 1 = bulk, 2 = finished (trivial right)
struct Key {
    uint8_t kind = 0, n = 0;
    uint16_t pos[2] = {0, 0};
    uint32_t tag[2] = {0, 0};
    bool operator<(Key const& o) const {
        if (kind != o.kind) return kind < o.kind;
        if (n != o.n) return n < o.n;
        for (int i = 0; i < n; ++i) {
            if (pos[i] != o.pos[i]) return pos[i] < o.pos[i];
            if (tag[i] != o.tag[i]) return tag[i] < o.tag[i];
        }
        return false;
    }
    bool operator==(Key const& o) const { return !(*this < o) && !(o < *this); }
};
Key key_left() { return Key{0, 0, {0, 0}, {0, 0}}; }
Key key_right() { return Key{2, 0, {0, 0}, {0, 0}}; }
Key key_ops(int n, const uint16_t* p, const uint32_t* t) {
    Key k;
    k.kind = 1;
    k.n = uint8_t(n);
    for (int i = 0; i < n; ++i) { k.pos[i] = p[i]; k.tag[i] = t[i]; }
    return k;
}

struct Cell {
    std::vector<std::pair<uint32_t, double>> terms;  // (tag, scale), same tags merged
    void add(uint32_t t, double s) {
        for (auto& x : terms)
            if (x.first == t) { x.second += s; return; }
        terms.push_back({t, s});
    }
};

// sector layout of a physical index given as a list of states with charges
struct PhysLayout {
    std::vector<Charge> charges;   // descending sectors
    std::vector<uint32_t> sizes;
    std::vector<int> state_sector, state_offset;  // per state
};

PhysLayout one_site_layout() {
    PhysLayout P;
    for (int s = 0; s < 4; ++s) {
        P.charges.push_back(SITE_Q[s]);
        P.sizes.push_back(1);
        P.state_sector.push_back(s);
        P.state_offset.push_back(0);
    }
    return P;
}

// fused two-site states (s1, s2) -> sector (charge sum, descending) and offset in the order of
// ProductBasis(phys, phys): s1 outer, s2 inner (indexing_stable.hpp:331-345)
PhysLayout two_site_layout() {
    PhysLayout P;
    std::map<Charge, uint32_t, std::greater<Charge>> sz;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) sz[SITE_Q[a] + SITE_Q[b]] += 1;
    for (auto& kv : sz) { P.charges.push_back(kv.first); P.sizes.push_back(kv.second); }
    std::map<Charge, int> fill;
    P.state_sector.resize(16);
    P.state_offset.resize(16);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            Charge c = SITE_Q[a] + SITE_Q[b];
            int sec = int(std::find(P.charges.begin(), P.charges.end(), c) - P.charges.begin());
            P.state_sector[a * 4 + b] = sec;
            P.state_offset[a * 4 + b] = fill[c]++;
        }
    return P;
}

// dense operator (dim x dim, M[out*dim+in]) -> site-op blocks appended to a FlatMPO
uint32_t append_op(FlatMPO& F, PhysLayout const& P, std::vector<double> const& M, int dim) {
    std::map<std::pair<int, int>, std::vector<std::tuple<int, int, double>>> blocks;  // (sec_in, sec_out)
    for (int o = 0; o < dim; ++o)
        for (int i = 0; i < dim; ++i) {
            double v = M[size_t(o) * dim + i];
            if (v == 0.0) continue;
            blocks[{P.state_sector[i], P.state_sector[o]}].push_back(
                std::make_tuple(P.state_offset[i], P.state_offset[o], v));
        }
    uint32_t op = uint32_t(F.op_spin.size());
    F.op_spin.push_back(0);
    // blocks ordered by descending (left, right) charges
    std::vector<std::pair<std::pair<int, int>, std::vector<std::tuple<int, int, double>>>> bl(blocks.begin(), blocks.end());
    std::sort(bl.begin(), bl.end(), [](auto const& a, auto const& b) { return a.first < b.first; });
    for (auto& b : bl) {
        Charge lc = P.charges[b.first.first], rc = P.charges[b.first.second];
        F.block_charges.insert(F.block_charges.end(), {lc.q[0], lc.q[1], rc.q[0], rc.q[1]});
        for (auto& e : b.second) {
            F.entry_rc.push_back(uint32_t(std::get<0>(e)));
            F.entry_rc.push_back(uint32_t(std::get<1>(e)));
            F.entry_coeff.push_back(std::get<2>(e));
        }
        F.block_entry_ptr.push_back(int64_t(F.entry_coeff.size()));
    }
    F.op_block_ptr.push_back(int64_t(F.block_charges.size() / 4));
    return op;
}

}  // namespace

struct QCBuilder {
    int L = 0;
    double ecore = 0;
    OpTable table;
    // per site: cells keyed by (k1, k2)
    std::vector<std::map<std::pair<Key, Key>, Cell>> prempo;
    std::vector<Mat4> site_terms;
    std::vector<char> has_site_term;
    std::vector<FlatMPO> site_mpo, two_mpo;
    std::vector<uint32_t> bond_dims;
    std::vector<std::vector<Key>> bond_keys;  // right-bond keys per site in index order
    std::vector<std::vector<uint32_t>> cell_terms;  // scratch
    uint32_t tag_id, tag_fill;

    void insert_fill(int from, int to, Key const& k, uint32_t fill) {
        for (int p = from; p < to; ++p) {
            auto& cell = prempo[p][{k, k}];
            if (cell.terms.empty()) cell.terms.push_back({fill, 1.0});
            else require(cell.terms.size() == 1 && cell.terms[0].first == fill, "qc mpo: conflicting fill operator");
        }
    }
    void attach(int p, Key const& k1, Key const& k2, uint32_t t) {
        auto& cell = prempo[p][{k1, k2}];
        if (cell.terms.empty()) cell.terms.push_back({t, 1.0});
    }
    void detach(int p, Key const& k1, Key const& k2, uint32_t t, double c) { prempo[p][{k1, k2}].add(t, c); }

    // parity of the product of the given tags
    bool odd(int n, const uint32_t* tags) {
        bool o = false;
        for (int i = 0; i < n; ++i) o ^= table.fermionic(tags[i]);
        return o;
    }

    void add_string(int n, const uint16_t* pos, const uint32_t* tag, double c, int& rightmost_left, int& leftmost_right) {
        // identity paths must reach every term's first site and leave from every term's last site
        // (TaggedMPOMaker::add_term, tagged_mpo_maker_optim.hpp:204-205)
        rightmost_left = std::max(rightmost_left, int(pos[0]));
        leftmost_right = std::min(leftmost_right, int(pos[n - 1]));
        if (n == 1) {
            Mat4 const& m = table.ops[tag[0]];
            for (int i = 0; i < 16; ++i) site_terms[pos[0]][i] += c * m[i];
            has_site_term[pos[0]] = 1;
            return;
        }
        Key TL = key_left(), TR = key_right();
        auto fill_for = [&](int first_pending) { return odd(n - first_pending, tag + first_pending) ? tag_fill : tag_id; };
        if (n == 2) {
            Key k = key_ops(1, pos + 1, tag + 1);
            detach(pos[0], TL, k, tag[0], c);
            insert_fill(pos[0] + 1, pos[1], k, fill_for(1));
            attach(pos[1], k, TR, tag[1]);
        } else if (n == 3) {
            Key k1 = key_ops(1, pos, tag);
            attach(pos[0], TL, k1, tag[0]);
            insert_fill(pos[0] + 1, pos[1], k1, fill_for(1));
            Key k2 = key_ops(1, pos + 2, tag + 2);
            detach(pos[1], k1, k2, tag[1], c);
            insert_fill(pos[1] + 1, pos[2], k2, fill_for(2));
            attach(pos[2], k2, TR, tag[2]);
        } else if (pos[1] + pos[2] < L - 1) {
            // gap between the 2nd and 3rd operator left of the centre: carry the applied pair
            // (the reference's add_4term keying)
            Key k1 = key_ops(1, pos, tag);
            attach(pos[0], TL, k1, tag[0]);
            insert_fill(pos[0] + 1, pos[1], k1, fill_for(1));
            Key k2 = key_ops(2, pos, tag);
            attach(pos[1], k1, k2, tag[1]);
            insert_fill(pos[1] + 1, pos[2], k2, fill_for(2));
            Key k3 = key_ops(1, pos + 3, tag + 3);
            detach(pos[2], k2, k3, tag[2], c);
            insert_fill(pos[2] + 1, pos[3], k3, fill_for(3));
            attach(pos[3], k3, TR, tag[3]);
        } else {
            // gap right of the centre: carry the pending pair instead, so that no bond holds more
            // than ~min(p, L-p)^2 pair states
            Key k1 = key_ops(1, pos, tag);
            attach(pos[0], TL, k1, tag[0]);
            insert_fill(pos[0] + 1, pos[1], k1, fill_for(1));
            Key k2 = key_ops(2, pos + 2, tag + 2);
            detach(pos[1], k1, k2, tag[1], c);
            insert_fill(pos[1] + 1, pos[2], k2, fill_for(2));
            Key k3 = key_ops(1, pos + 3, tag + 3);
            attach(pos[2], k2, k3, tag[2]);
            insert_fill(pos[2] + 1, pos[3], k3, fill_for(3));
            attach(pos[3], k3, TR, tag[3]);
        }
    }

    void build(const double* h1, const double* eri, double cutoff) {
        prempo.assign(L, {});
        site_terms.assign(L, Mat4{});
        has_site_term.assign(L, 0);
        tag_id = table.tag(op_ident());
        tag_fill = table.tag(op_fill());
        Mat4 cd[2] = {op_cdag(0), op_cdag(1)}, cc[2] = {op_c(0), op_c(1)};
        Mat4 F = op_fill(), I = op_ident();
        // operator strings in Jordan-Wigner form, merged by (site, local operator) content
        struct SKey {
            uint8_t n;
            uint16_t pos[4];
            uint32_t tag[4];
            bool operator==(SKey const& o) const {
                if (n != o.n) return false;
                for (int i = 0; i < n; ++i)
                    if (pos[i] != o.pos[i] || tag[i] != o.tag[i]) return false;
                return true;
            }
        };
        struct SHash {
            size_t operator()(SKey const& k) const {
                size_t h = k.n;
                for (int i = 0; i < k.n; ++i) h = h * 1000003u ^ (size_t(k.pos[i]) << 20 ^ k.tag[i]);
                return h;
            }
        };
        std::unordered_map<SKey, double, SHash> strings;
        // ops: (site, is_creator, spin) in product order
        auto add_product = [&](int n, const int* site, const int* cre, const int* spin, double coeff) {
            if (std::fabs(coeff) <= cutoff) return;
            int lo = site[0], hi = site[0];
            for (int i = 1; i < n; ++i) { lo = std::min(lo, site[i]); hi = std::max(hi, site[i]); }
            SKey k{};
            k.n = 0;
            for (int q = lo; q <= hi; ++q) {
                bool on = false;
                for (int i = 0; i < n; ++i) on |= site[i] == q;
                if (!on) continue;
                Mat4 m = I;
                for (int i = 0; i < n; ++i) {
                    if (site[i] > q) m = mul(m, F);
                    else if (site[i] == q) m = mul(m, cre[i] ? cd[spin[i]] : cc[spin[i]]);
                }
                if (is_zero(m)) return;
                // canonical sign (first non-zero entry positive); the sign moves to the coefficient
                for (double v : m)
                    if (v != 0.0) {
                        if (v < 0) {
                            for (auto& x : m) x = -x;
                            coeff = -coeff;
                        }
                        break;
                    }
                k.pos[k.n] = uint16_t(q);
                k.tag[k.n] = table.tag(m);
                ++k.n;
            }
            strings[k] += coeff;
        };
        for (int p = 0; p < L; ++p)
            for (int q = 0; q < L; ++q) {
                double h = h1[size_t(p) * L + q];
                for (int s = 0; s < 2; ++s) {
                    int site[2] = {p, q}, cre[2] = {1, 0}, spin[2] = {s, s};
                    add_product(2, site, cre, spin, h);
                }
            }
        for (int p = 0; p < L; ++p)
            for (int q = 0; q < L; ++q)
                for (int r = 0; r < L; ++r)
                    for (int s = 0; s < L; ++s) {
                        double v = 0.5 * eri[((size_t(p) * L + q) * L + r) * L + s];
                        if (std::fabs(v) <= cutoff) continue;
                        for (int sg = 0; sg < 2; ++sg)
                            for (int tg = 0; tg < 2; ++tg) {
                                if ((p == r && sg == tg) || (q == s && sg == tg)) continue;
                                int site[4] = {p, r, s, q}, cre[4] = {1, 1, 0, 0}, spin[4] = {sg, tg, tg, sg};
                                add_product(4, site, cre, spin, v);
                            }
                    }
        int rightmost_left = 0, leftmost_right = L - 1;
        // deterministic order of the merged strings
        std::vector<std::pair<SKey, double>> ordered(strings.begin(), strings.end());
        std::sort(ordered.begin(), ordered.end(), [](auto const& a, auto const& b) {
            if (a.first.n != b.first.n) return a.first.n < b.first.n;
            for (int i = 0; i < a.first.n; ++i) {
                if (a.first.pos[i] != b.first.pos[i]) return a.first.pos[i] < b.first.pos[i];
                if (a.first.tag[i] != b.first.tag[i]) return a.first.tag[i] < b.first.tag[i];
            }
            return false;
        });
        for (auto const& s : ordered) {
            if (std::fabs(s.second) <= cutoff) continue;
            add_string(s.first.n, s.first.pos, s.first.tag, s.second, rightmost_left, leftmost_right);
        }
        // site terms and identity paths (TaggedMPOMaker::finalize)
        Key TL = key_left(), TR = key_right();
        for (int p = 0; p < L; ++p)
            if (has_site_term[p] && !is_zero(site_terms[p])) {
                uint32_t t = table.tag(site_terms[p]);
                prempo[p][{TL, TR}].add(t, 1.0);
            }
        for (int p = 0; p < rightmost_left; ++p) attach(p, TL, TL, tag_id);
        insert_fill(leftmost_right + 1, L, TR, tag_id);
        assemble();
    }

    // Hermitian partner of a key: operators transposed (abelian phase 1)
    Key conj_key(Key k) {
        for (int i = 0; i < k.n; ++i) k.tag[i] = table.tag(transpose(table.ops[k.tag[i]]));
        return k;
    }
    bool self_adjoint(Key const& k) {
        if (k.kind != 1) return false;
        for (int i = 0; i < k.n; ++i)
            if (!(transpose(table.ops[k.tag[i]]) == table.ops[k.tag[i]])) return false;
        return true;
    }

    void assemble() {
        PhysLayout P1 = one_site_layout();
        site_mpo.assign(L, FlatMPO());
        bond_dims.assign(L + 1, 1);
        bond_keys.assign(L + 1, {});
        std::map<Key, uint32_t> left;
        left[key_left()] = 0;
        bond_keys[0] = {key_left()};
        std::vector<uint32_t> lherm{MQ_NO_PARTNER};
        std::vector<int32_t> lphase{1};
        for (int p = 0; p < L; ++p) {
            std::map<Key, uint32_t> right;
            std::vector<Key> rk;
            uint32_t r = 2;
            // right keys in sorted order: unfinished -> 0, finished -> 1, bulk from 2
            std::vector<Key> k2s;
            for (auto const& kv : prempo[p]) k2s.push_back(kv.first.second);
            std::sort(k2s.begin(), k2s.end());
            k2s.erase(std::unique(k2s.begin(), k2s.end()), k2s.end());
            for (auto const& k : k2s) {
                if (p == L - 1) { right[k] = 0; continue; }
                if (k.kind == 0) right[k] = 0;
                else if (k.kind == 2) right[k] = 1;
                else right[k] = r++;
            }
            uint32_t rdim = (p == L - 1) ? 1 : std::max<uint32_t>(r, 2);
            FlatMPO& F = site_mpo[p];
            F.row_dim = uint32_t(bond_dims[p]);
            F.col_dim = rdim;
            // terms sorted by (b1, b2)
            std::map<std::pair<uint32_t, uint32_t>, Cell const*> cells;
            for (auto const& kv : prempo[p]) {
                auto li = left.find(kv.first.first);
                require(li != left.end(), "qc mpo: left key not found");
                cells[{li->second, right[kv.first.second]}] = &kv.second;
            }
            std::map<uint32_t, uint32_t> tag_op;
            for (auto const& c : cells)
                for (auto const& t : c.second->terms) {
                    if (t.second == 0.0) continue;
                    if (!tag_op.count(t.first)) {
                        std::vector<double> M(table.ops[t.first].begin(), table.ops[t.first].end());
                        tag_op[t.first] = append_op(F, P1, M, 4);
                    }
                    F.term_b.push_back(c.first.first);
                    F.term_b.push_back(c.first.second);
                    F.term_op.push_back(tag_op[t.first]);
                    F.term_scale.push_back(t.second);
                }
            F.lspin.assign(F.row_dim, 0);
            F.rspin.assign(F.col_dim, 0);
            F.lherm = lherm;
            F.lphase = lphase;
            F.lherm.resize(F.row_dim, MQ_NO_PARTNER);
            F.lphase.resize(F.row_dim, 1);
            std::vector<uint32_t> rherm(F.col_dim, MQ_NO_PARTNER);
            std::vector<int32_t> rphase(F.col_dim, 1);
            if (p < L - 1)
                for (auto const& kv : right) {
                    if (kv.first.kind != 1) continue;
                    if (self_adjoint(kv.first)) { rherm[kv.second] = kv.second; continue; }
                    auto it = right.find(conj_key(kv.first));
                    if (it != right.end() && it->second != kv.second) rherm[kv.second] = it->second;
                }
            F.rherm = rherm;
            F.rphase = rphase;
            bond_dims[p + 1] = F.col_dim;
            left = right;
            lherm = rherm;
            lphase = rphase;
        }
        prempo.clear();
    }

    // two-site MPO of sites (p, p+1): cells (b0, b2) = sum_b1 W_p(b0,b1) (x) W_p+1(b1,b2)
    // (reference: make_twosite_mpo, mp_tensors/ts_ops.hpp:96-168)
    FlatMPO twosite(int p) {
        FlatMPO const& A = site_mpo[p];
        FlatMPO const& B = site_mpo[p + 1];
        PhysLayout P2 = two_site_layout();
        FlatMPO T;
        T.row_dim = A.row_dim;
        T.col_dim = B.col_dim;
        // dense 4x4 of every op of A and B
        auto dense = [](FlatMPO const& F, uint32_t op) {
            Mat4 m{};
            for (int64_t bl = F.op_block_ptr[op]; bl < F.op_block_ptr[op + 1]; ++bl) {
                Charge lc(F.block_charges[4 * bl], F.block_charges[4 * bl + 1]);
                Charge rc(F.block_charges[4 * bl + 2], F.block_charges[4 * bl + 3]);
                int in = 0, out = 0;
                for (int s = 0; s < 4; ++s) { if (SITE_Q[s] == lc) in = s; if (SITE_Q[s] == rc) out = s; }
                for (int64_t e = F.block_entry_ptr[bl]; e < F.block_entry_ptr[bl + 1]; ++e) m[out * 4 + in] += F.entry_coeff[e];
            }
            return m;
        };
        std::vector<Mat4> Ad(A.op_spin.size()), Bd(B.op_spin.size());
        for (size_t o = 0; o < Ad.size(); ++o) Ad[o] = dense(A, uint32_t(o));
        for (size_t o = 0; o < Bd.size(); ++o) Bd[o] = dense(B, uint32_t(o));
        // rows of B by b1
        std::vector<std::vector<size_t>> brow(B.row_dim);
        for (size_t t = 0; t < B.term_op.size(); ++t) brow[B.term_b[2 * t]].push_back(t);
        std::map<std::pair<uint32_t, uint32_t>, std::map<std::pair<uint32_t, uint32_t>, double>> cells;
        for (size_t ta = 0; ta < A.term_op.size(); ++ta) {
            uint32_t b0 = A.term_b[2 * ta], b1 = A.term_b[2 * ta + 1];
            for (size_t tb : brow[b1]) {
                uint32_t b2 = B.term_b[2 * tb + 1];
                cells[{b0, b2}][{A.term_op[ta], B.term_op[tb]}] += A.term_scale[ta] * B.term_scale[tb];
            }
        }
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> kron_op;
        for (auto const& c : cells)
            for (auto const& t : c.second) {
                if (t.second == 0.0) continue;
                auto it = kron_op.find(t.first);
                if (it == kron_op.end()) {
                    std::vector<double> K(256, 0.0);
                    Mat4 const& a = Ad[t.first.first];
                    Mat4 const& b = Bd[t.first.second];
                    for (int o1 = 0; o1 < 4; ++o1)
                        for (int i1 = 0; i1 < 4; ++i1) {
                            double va = a[o1 * 4 + i1];
                            if (va == 0.0) continue;
                            for (int o2 = 0; o2 < 4; ++o2)
                                for (int i2 = 0; i2 < 4; ++i2) K[size_t(o1 * 4 + o2) * 16 + (i1 * 4 + i2)] += va * b[o2 * 4 + i2];
                        }
                    uint32_t op = append_op(T, P2, K, 16);
                    it = kron_op.emplace(t.first, op).first;
                }
                T.term_b.push_back(c.first.first);
                T.term_b.push_back(c.first.second);
                T.term_op.push_back(it->second);
                T.term_scale.push_back(t.second);
            }
        T.lspin = A.lspin;
        T.rspin = B.rspin;
        T.lherm = A.lherm;
        T.lphase = A.lphase;
        T.rherm = B.rherm;
        T.rphase = B.rphase;
        return T;
    }
};

}  // namespace mq

struct mq_qc_s {
    int sym = MQ_SYM_TWOU1;
    mq::QCBuilder b;                                          // TwoU1
    std::unique_ptr<mq::SU2QC, mq::SU2QCDeleter> su2;         // SU2U1
    std::vector<std::unique_ptr<mq::FlatMPO>> two;
    std::vector<int32_t> phys_charges[2];
    std::vector<uint32_t> phys_sizes[2];
    std::string err;
    int length() const { return sym == MQ_SYM_SU2U1 ? mq::su2qc_length(*su2) : b.L; }
};

extern "C" {

int mq_qc_build(int symmetry, int L, const double* h1, const double* eri, double ecore, double cutoff, mq_qc_t* out) {
    return mq_qc_build_ex(symmetry, L, h1, eri, ecore, cutoff, 0, out);
}

int mq_qc_build_ex(int symmetry, int L, const double* h1, const double* eri, double ecore, double cutoff, int flags,
                   mq_qc_t* out) {
    if (!out || L < 2 || !h1 || !eri || (flags & ~MQ_QC_GRADIENT)) return MQ_ERR_INVALID;
    if (symmetry != MQ_SYM_TWOU1 && symmetry != MQ_SYM_SU2U1) return MQ_ERR_UNSUPPORTED;
    if ((flags & MQ_QC_GRADIENT) && symmetry != MQ_SYM_SU2U1) {
        fprintf(stderr, "mq_qc_build_ex: the gradient layout is built for SU2U1 chains only\n");
        return MQ_ERR_UNSUPPORTED;
    }
    try {
        std::unique_ptr<mq_qc_s> q(new mq_qc_s());
        q->sym = symmetry;
        q->b.ecore = ecore;
        if (symmetry == MQ_SYM_SU2U1) {
            q->b.L = L;
            q->su2 = mq::su2qc_build(L, h1, eri, cutoff, (flags & MQ_QC_GRADIENT) != 0);
            for (int k = 0; k < 2; ++k) mq::su2qc_phys(k, q->phys_charges[k], q->phys_sizes[k]);
        } else {
            q->b.L = L;
            q->b.build(h1, eri, cutoff);
            q->two.resize(L - 1);
            for (int k = 0; k < 2; ++k) {
                mq::PhysLayout P = k ? mq::two_site_layout() : mq::one_site_layout();
                for (size_t i = 0; i < P.charges.size(); ++i) {
                    q->phys_charges[k].push_back(P.charges[i].q[0]);
                    q->phys_charges[k].push_back(P.charges[i].q[1]);
                    q->phys_sizes[k].push_back(P.sizes[i]);
                }
            }
        }
        *out = q.release();
        return MQ_OK;
    } catch (mq::Error const& e) {
        fprintf(stderr, "mq_qc_build: %s\n", e.what());
        return e.code;
    } catch (std::exception const& e) {
        fprintf(stderr, "mq_qc_build: %s\n", e.what());
        return MQ_ERR_INVALID;
    }
}

int mq_qc_info(mq_qc_t q, int* L, uint32_t* bond_dims, double* ecore) {
    if (!q) return MQ_ERR_INVALID;
    if (L) *L = q->length();
    if (bond_dims) {
        auto const& bd = q->sym == MQ_SYM_SU2U1 ? mq::su2qc_bond_dims(*q->su2) : q->b.bond_dims;
        std::copy(bd.begin(), bd.end(), bond_dims);
    }
    if (ecore) *ecore = q->b.ecore;
    return MQ_OK;
}

int mq_qc_site_mpo(mq_qc_t q, int site, int twosite, mq_mpo_desc* out) {
    if (!q || !out || site < 0 || site >= q->length() || (twosite && site >= q->length() - 1)) return MQ_ERR_INVALID;
    try {
        if (q->sym == MQ_SYM_SU2U1) {
            (twosite ? mq::su2qc_twosite(*q->su2, site) : mq::su2qc_site(*q->su2, site)).fill_desc(*out);
        } else if (!twosite) {
            q->b.site_mpo[site].fill_desc(*out);
        } else {
            if (!q->two[site]) q->two[site].reset(new mq::FlatMPO(q->b.twosite(site)));
            q->two[site]->fill_desc(*out);
        }
        return MQ_OK;
    } catch (std::exception const& e) {
        fprintf(stderr, "mq_qc_site_mpo: %s\n", e.what());
        return MQ_ERR_INVALID;
    }
}

int mq_qc_phys(mq_qc_t q, int twosite, mq_index_desc* out) {
    if (!q || !out) return MQ_ERR_INVALID;
    int k = twosite ? 1 : 0;
    out->n = int32_t(q->phys_sizes[k].size());
    out->charges = q->phys_charges[k].data();
    out->sizes = q->phys_sizes[k].data();
    return MQ_OK;
}

int mq_qc_integral_gradient(mq_qc_t q, const double* const* term_grads, double* dh, double* deri) {
    if (!q || !term_grads || !dh || !deri) return MQ_ERR_INVALID;
    if (q->sym != MQ_SYM_SU2U1) return MQ_ERR_UNSUPPORTED;
    try {
        const int L = q->length();
        std::vector<const double*> e(term_grads, term_grads + L);
        for (const double* x : e) mq::require(x != nullptr, "mq_qc_integral_gradient: null term-gradient array");
        mq::su2qc_integral_gradient(*q->su2, e, dh, deri);
        return MQ_OK;
    } catch (mq::Error const& e) {
        fprintf(stderr, "mq_qc_integral_gradient: %s\n", e.what());
        return e.code;
    } catch (std::exception const& e) {
        fprintf(stderr, "mq_qc_integral_gradient: %s\n", e.what());
        return MQ_ERR_INVALID;
    }
}

int mq_qc_destroy(mq_qc_t q) {
    delete q;
    return MQ_OK;
}

}  // extern "C"
