// Tile passes compiled at run time (hiprtc) for their exact gate structure.
//
// The interpreted tile kernel (kq_tile.hip) dispatches every gate through a switch over ~150 variants;
// each dispatch costs register moves and branches comparable to the arithmetic of a cheap gate.  A
// generated pass kernel has the pass's structure (layouts, remaps, gate variants, register positions,
// controls, merged diagonal runs) as straight-line code with literal constants, so swaps become register
// renaming and no dispatch remains.  Only the coefficients are data: they are read from a per-pass buffer
// of doubles, so passes whose structure repeats with other angles (optimiser iterations) reuse the
// compiled kernel.  Kernels are cached by their generated source; missing ones compile in parallel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <bit>
#include <atomic>
#include <cmath>
#include <complex>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <functional>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kq_internal.hpp"

namespace kq
{
namespace
{

// the device templates, embedded at build time from kq_tile_device.hpp
const char* const DEVICE_HEADER =
#include "kq_tile_device.inc"
    ;

struct JitCache
{
    std::mutex mu;
    std::unordered_map<std::string, hipFunction_t> kernels;
    std::vector<hipModule_t> modules;
    JitStats stats {};
    int mode = 1;            // 0 off, 1 large states, 2 every tile pass
    int available = -1;      // hiprtc usable (-1 unknown)
    JitTuning tuning {};
};

JitCache& cache()
{
    static JitCache c;
    return c;
}

std::string hex64(uint64_t x)
{
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxull", static_cast<unsigned long long>(x));
    return b;
}

std::string hex32(uint32_t x)
{
    char b[24];
    std::snprintf(b, sizeof b, "0x%xu", x);
    return b;
}

// block index deposited into the free qubits, as runs of consecutive bits
std::string deposit_expr(const std::vector<uint8_t>& freepos)
{
    std::string e;
    size_t i = 0, src = 0;
    while (i < freepos.size()) {
        size_t j = i;
        while (j + 1 < freepos.size() && freepos[j + 1] == freepos[j] + 1) {
            ++j;
        }
        const size_t len = j - i + 1;
        if (!e.empty()) {
            e += " | ";
        }
        e += "(((blk >> " + std::to_string(src) + ") & " + hex64((len >= 64) ? ~0ull : ((1ull << len) - 1)) + ") << "
           + std::to_string(freepos[i]) + ")";
        src += len;
        i = j + 1;
    }
    return e.empty() ? std::string {"0ull"} : e;
}

std::string lds_thread_expr(const TileLayout& L, int nbits)
{
    std::string e = "0u";
    for (int j = 0; j < nbits; ++j) {
        e += " ^ ((tid & " + hex32(1u << j) + ") ? " + hex32(L.lds_thr[j]) + " : 0u)";
    }
    return e;
}

// LDS bank rules of gfx950 for one access width (measured with conflict counters): a write serves
// the lanes that differ in thread bits below `write_lanes` in one cycle and conflicts unless their
// entry indices differ in bits below `write_lanes` (entry = one access); a read serves the lanes
// below `read_lanes` and conflicts unless their entries differ in the low `read_lanes` bits, where
// for 16-byte reads the top one of those bits is entry bit 3 ^ entry bit 4.
struct BankRule
{
    int write_lanes;
    int read_lanes;
    bool read_hash34;
};
constexpr BankRule BANKS_8B {4, 5, false};

// The LDS entries of one remap.  The tile is stored at entry F(l) of local index l, F(l) = l ^ G(l)
// linear and unit lower triangular (local bit i may flip entry bits below min(i, read_lanes)), so F
// is invertible whatever G is.  G is searched so that the writing layout's low lanes (each lane bit a
// local bit, plus the register bit of any pending controlled X conditioned on that lane bit alone,
// which the write address folds in) and the reading layout's low lanes meet the bank rule.  Without
// a solution in a bounded search the layouts are returned unchanged.  Offsets are returned in bytes
// of `entry_bytes` per entry.
std::pair<TileLayout, TileLayout> bank_spread(const TileLayout& A, const TileLayout& B, int k, int nthr, size_t ab,
                                              size_t entry_bytes, const BankRule& rule,
                                              const std::vector<std::pair<int, int>>& lane_pending)
{
    const int wl = rule.write_lanes, rl = rule.read_lanes;
    const auto local = [&](uint32_t bytes) { return lds_swizzle(bytes / static_cast<uint32_t>(ab)); };
    const auto rescale = [&](const TileLayout& L) {
        TileLayout R = L;
        for (int r = 0; r < 16; ++r) {
            R.lds_slot[r] = local(L.lds_slot[r]) * static_cast<uint32_t>(entry_bytes);
        }
        for (int j = 0; j < nthr; ++j) {
            R.lds_thr[j] = local(L.lds_thr[j]) * static_cast<uint32_t>(entry_bytes);
        }
        return R;
    };
    if (nthr < rl || k <= rl || k > 12) {
        return {rescale(A), rescale(B)};
    }
    uint32_t wv[8], rv[8];
    for (int j = 0; j < rl; ++j) {
        const uint32_t a = local(A.lds_thr[j]), b = local(B.lds_thr[j]);
        if (std::popcount(a) != 1 || std::popcount(b) != 1) {
            return {rescale(A), rescale(B)};
        }
        if (j < wl) {
            wv[j] = a;
        }
        rv[j] = b;
    }
    for (const auto& [lane, P] : lane_pending) {
        if (lane < wl) {
            wv[lane] ^= local(A.lds_slot[1u << P]);
        }
    }
    uint32_t g[12] = {};
    bool assigned[12] = {};
    const int lo = std::min(wl, rl);
    for (int i = 0; i < lo; ++i) {
        assigned[i] = true;
    }
    const auto F = [&](uint32_t l) {
        uint32_t e = l;
        for (int i = lo; i < k; ++i) {
            if ((l >> i) & 1u) {
                e ^= g[i];
            }
        }
        return e;
    };
    const auto wbank = [&](uint32_t l) { return F(l) & ((1u << wl) - 1); };
    const auto rbank = [&](uint32_t l) {
        const uint32_t e = F(l);
        if (rule.read_hash34) {
            return (e & 7u) | ((((e >> 3) ^ (e >> 4)) & 1u) << 3);
        }
        return e & ((1u << rl) - 1);
    };
    const auto ready = [&](uint32_t l) {
        for (int i = lo; i < k; ++i) {
            if (((l >> i) & 1u) && !assigned[i]) {
                return false;
            }
        }
        return true;
    };
    // the assigned vectors of a set must be independent (every non-empty subset XOR non-zero)
    const auto ok = [&](const uint32_t* vecs, int n, bool write) {
        uint32_t bank[8];
        int m = 0;
        for (int j = 0; j < n; ++j) {
            if (ready(vecs[j])) {
                bank[m++] = write ? wbank(vecs[j]) : rbank(vecs[j]);
            }
        }
        for (uint32_t sub = 1; sub < (1u << m); ++sub) {
            uint32_t x = 0;
            for (int j = 0; j < m; ++j) {
                if ((sub >> j) & 1u) {
                    x ^= bank[j];
                }
            }
            if (x == 0) {
                return false;
            }
        }
        return true;
    };
    std::vector<int> free_bits;
    const auto collect = [&](uint32_t l) {
        for (int i = lo; i < k; ++i) {
            if (((l >> i) & 1u) && std::find(free_bits.begin(), free_bits.end(), i) == free_bits.end()) {
                free_bits.push_back(i);
            }
        }
    };
    for (int j = 0; j < rl; ++j) {
        collect(rv[j]);
    }
    for (int j = 0; j < wl; ++j) {
        collect(wv[j]);
    }
    size_t budget = 200000;
    const std::function<bool(size_t)> search = [&](size_t i) {
        if (i == free_bits.size()) {
            return true;
        }
        const int b = free_bits[i];
        const uint32_t n_values = 1u << std::min(b, rl);
        for (uint32_t v = 0; v < n_values && budget > 0; ++v, --budget) {
            g[b] = v;
            assigned[b] = true;
            if (ok(wv, wl, true) && ok(rv, rl, false) && search(i + 1)) {
                return true;
            }
        }
        assigned[b] = false;
        g[b] = 0;
        return false;
    };
    if (!ok(wv, wl, true) || !ok(rv, rl, false) || !search(0)) {
        return {rescale(A), rescale(B)};
    }
    TileLayout A2 = A, B2 = B;
    const uint32_t eb = static_cast<uint32_t>(entry_bytes);
    for (int r = 0; r < 16; ++r) {
        A2.lds_slot[r] = F(local(A.lds_slot[r])) * eb;
        B2.lds_slot[r] = F(local(B.lds_slot[r])) * eb;
    }
    for (int j = 0; j < nthr; ++j) {
        A2.lds_thr[j] = F(local(A.lds_thr[j])) * eb;
        B2.lds_thr[j] = F(local(B.lds_thr[j])) * eb;
    }
    return {A2, B2};
}

std::string hbm_thread_expr(const TileLayout& L, int nbits)
{
    std::string e = "0ull";
    for (int j = 0; j < nbits; ++j) {
        e += " | ((tid & " + hex32(1u << j) + ") ? " + hex64(L.hbm_thr[j]) + " : 0ull)";
    }
    return e;
}

std::string cond_expr(uint64_t glb_ctrl, uint32_t lane_ctrl)
{
    std::string c = "true";
    if (glb_ctrl) {
        c += " && ((base & " + hex64(glb_ctrl) + ") == " + hex64(glb_ctrl) + ")";
    }
    if (lane_ctrl) {
        c += " && ((tid & " + hex32(lane_ctrl) + ") == " + hex32(lane_ctrl) + ")";
    }
    return c;
}

std::string sel_expr(uint32_t lane_sel, uint64_t glb_sel)
{
    if (lane_sel) {
        return "((tid & " + hex32(lane_sel) + ") != 0u)";
    }
    return "((base & " + hex64(glb_sel) + ") != 0ull)";
}

// the 8 coefficients of a record as doubles (complex64 records hold 8 packed floats in m[0..3])
std::array<double, 8> record_values(const TileGate& g, bool c64)
{
    std::array<double, 8> m {};
    if (!c64) {
        std::copy(g.m, g.m + 8, m.begin());
        return m;
    }
    float f[8];
    std::memcpy(f, g.m, sizeof f);
    std::copy(f, f + 8, m.begin());
    return m;
}

// Emits an uncontrolled one-qubit gate with its pivot entry divided out (orth / axis / dense1 / a
// d0-normalised diagonal) and multiplies the pivot into the pass factor; false if the record has
// another shape.  The normalised values are data; which entry is the pivot is structure.
// `flip`: a condition (thread / outside bits) under which controlled X gates on the gate's register
// bit are still pending; the gate is then emitted conjugated by X under it (X G X), which leaves the
// X gates pending.  Only forms closed under that conjugation take it: an axis rotation is its own
// conjugate, a rotation with the diagonal pivot [[1, t], [-t, 1]] becomes [[1, -t], [t, 1]].
bool emit_normalised(std::ostringstream& o, const TileGate& g, bool c64, std::vector<double>& coefs,
                     std::complex<double>& factor, bool& factor_complex, const std::string& flip = {},
                     const std::vector<std::string>& vecs = {"v"}, double floor_abs = 0.0)
{
    if (g.code >= TV_UNIFORM_BASE || g.lane_ctrl != 0 || g.glb_ctrl != 0 || g.code % 5 != 0) {
        return false;
    }
    const int kind = static_cast<int>(g.code / 20);
    const int P = static_cast<int>((g.code / 5) % 4);
    const auto m = record_values(g, c64);
    const auto val = [&](double x) {
        coefs.push_back(x);
        return "static_cast<R>(cf[" + std::to_string(coefs.size() - 1) + "])";
    };
    // The diagonal entry is the pivot unless it is below 1e-6 of the off-diagonal one: c * (a + t b)
    // rounds like c a + s b for any finite t, so only a (nearly) vanishing pivot needs the other
    // form, and the kernel's structure stays the same for almost every angle.
    const auto diag_pivot_of = [](double diag, double off) { return std::abs(diag) >= 1.0e-6 * std::abs(off); };
    // amplitudes grow by 1 / |factor| inside the pass: keep that far from overflow
    const double floor = floor_abs > 0.0 ? floor_abs : (c64 ? 1.0e-20 : 1.0e-200);
    // the statement for every vector the pass carries (the adjoint's psi and lambda take the same gate)
    const auto each = [&](const std::string& head, const std::string& args) {
        for (const std::string& vec : vecs) {
            o << "  " << head << "(" << vec << ", " << args << ");\n";
        }
    };
    const auto fits = [&](std::complex<double> pivot) { return std::abs(factor * pivot) >= floor; };
    if (kind == TK_REAL) {
        const bool rotation = m[6] == m[0] && m[4] == -m[2];
        const bool reflection = m[6] == -m[0] && m[4] == m[2];
        if (!rotation && !reflection) {
            return false;
        }
        const bool diag_pivot = diag_pivot_of(m[0], m[2]);
        const double pivot = diag_pivot ? m[0] : m[2];
        if (pivot == 0.0 || !fits(pivot)) {
            return false;
        }
        const int form = (diag_pivot ? 0 : 2) + (rotation ? 0 : 1);
        if (!flip.empty() && form != 0) {
            return false;
        }
        const std::string t = val((diag_pivot ? m[2] : m[0]) / pivot);
        each("orth<R, " + std::to_string(P) + ", " + std::to_string(form) + ">",
             flip.empty() ? t : "((" + flip + ") ? -" + t + " : " + t + ")");
        factor *= pivot;
        return true;
    }
    if (kind == TK_AXIS) {
        if (m[6] != m[0] || m[5] != m[3]) {
            return false;
        }
        const bool diag_pivot = diag_pivot_of(m[0], m[3]);
        const double pivot = diag_pivot ? m[0] : m[3];
        if (pivot == 0.0 || !fits(pivot)) {
            return false;
        }
        each("axis<R, " + std::to_string(P) + ", " + std::to_string(diag_pivot ? 0 : 1) + ">", val((diag_pivot ? m[3] : m[0]) / pivot));
        factor *= pivot;
        return true;
    }
    if (!flip.empty() && kind != TK_DENSE) {
        return false;
    }
    if (kind == TK_DIAG) {
        const std::complex<double> d0 {m[0], m[1]};
        if (std::abs(d0) < 0.5 || !fits(d0)) {
            return false;
        }
        const std::complex<double> d = std::complex<double> {m[6], m[7]} / d0;
        const auto dr = val(d.real());
        const auto di = val(d.imag());
        o << "  { const Coef<R> m {{0, 0, 0, 0, 0, 0, " << dr << ", " << di << "}};";
        for (const std::string& vec : vecs) {
            o << " apply<R, K_DIAG1, " << P << ", 0>(" << vec << ", m);";
        }
        o << " }\n";
        factor *= d0;
        factor_complex = true;
        return true;
    }
    if (kind == TK_DENSE) {
        const std::complex<double> m00 {m[0], m[1]}, m01 {m[2], m[3]}, m10 {m[4], m[5]}, m11 {m[6], m[7]};
        const bool diag_pivot = diag_pivot_of(std::abs(m00), std::abs(m01));
        const std::complex<double> pivot = diag_pivot ? m00 : m01;
        if (std::abs(pivot) == 0.0 || !fits(pivot)) {
            return false;
        }
        const std::complex<double> x = (diag_pivot ? m01 : m00) / pivot, y = m10 / pivot, z = m11 / pivot;
        // A complex matrix whose normalised form is an axis rotation [[1, i t], [i t, 1]] (SX, SXDAG) or
        // a real rotation / reflection takes the cheaper form: 4 FP64 operations per pair instead of 12.
        // Entries within a few ulps of the form's zeros count as zero (a change below 1e-15 per gate).
        constexpr double ULPS = 4.0e-16;
        const auto near = [](double a, double b) { return std::abs(a - b) <= ULPS * std::max(1.0, std::abs(b)); };
        if (diag_pivot && near(z.imag(), 0.0)) {
            if (near(z.real(), 1.0) && near(x.real(), 0.0) && near(y.real(), 0.0) && near(y.imag(), x.imag())) {
                each("axis<R, " + std::to_string(P) + ", 0>", val(x.imag()));
                factor *= pivot;
                factor_complex = true;
                return true;
            }
            if (near(x.imag(), 0.0) && near(y.imag(), 0.0)) {
                const bool rotation = near(z.real(), 1.0) && near(y.real(), -x.real());
                const bool reflection = near(z.real(), -1.0) && near(y.real(), x.real());
                if ((rotation || reflection) && (flip.empty() || rotation)) {
                    const std::string t = val(x.real());
                    each("orth<R, " + std::to_string(P) + ", " + std::to_string(rotation ? 0 : 1) + ">",
                         flip.empty() ? t : "((" + flip + ") ? -" + t + " : " + t + ")");
                    factor *= pivot;
                    factor_complex = true;
                    return true;
                }
            }
        }
        if (!flip.empty()) {
            return false;
        }
        const auto xr = val(x.real()), xi = val(x.imag()), yr = val(y.real()), yi = val(y.imag());
        const auto zr = val(z.real()), zi = val(z.imag());
        each("dense1<R, " + std::to_string(P) + ", " + std::to_string(diag_pivot ? 0 : 1) + ">",
             xr + ", " + xi + ", " + yr + ", " + yi + ", " + zr + ", " + zi);
        factor *= pivot;
        factor_complex = true;
        return true;
    }
    return false;
}

// the two diagonal entries of a merged-run term (complex64 terms hold packed floats in d[0..1])
std::array<std::complex<double>, 2> term_values(const TileTerm& t, bool c64)
{
    if (!c64) {
        return {std::complex<double> {t.d[0], t.d[1]}, std::complex<double> {t.d[2], t.d[3]}};
    }
    float f[4];
    std::memcpy(f, t.d, sizeof f);
    return {std::complex<double> {f[0], f[1]}, std::complex<double> {f[2], f[3]}};
}

// diag(d0, d1) on register bit key / 5 with register control key % 5 (d0_one: d0 is exactly 1).  Under
// `flip` (controlled X gates still pending on that bit, key % 5 == 0) the lanes where it holds apply
// diag(d1, d0), and the X gates stay pending.
void emit_diag_apply(std::ostringstream& o, uint32_t key, const std::string& d0r, const std::string& d0i,
                     const std::string& d1r, const std::string& d1i, bool d0_one, const std::string& flip)
{
    const std::string m = "{{" + d0r + ", " + d0i + ", 0, 0, 0, 0, " + d1r + ", " + d1i + "}}";
    if (flip.empty()) {
        o << "  { const Coef<R> m " << m << "; apply<R, " << (d0_one ? "K_DIAG1" : "K_DIAG") << ", " << key / 5 << ", "
          << key % 5 << ">(v, m); }\n";
    }
    else {
        o << "  { const Coef<R> m = flip_coef<R>(Coef<R> " << m << ", " << flip << "); apply<R, K_DIAG, " << key / 5
          << ", 0>(v, m); }\n";
    }
}

// The factor of a merged run of pure-phase terms with d0 = 1 (or, for a run on a thread or outside
// bit, a factor selected by that bit) from tables built on the host: the product over the terms
// conditioned on thread bits only is one entry of a table indexed by the thread index; the terms
// conditioned on outside bits only are grouped by the bits they read (at most 6 per group) into
// tables indexed by those bits of the tile's position; terms reading both kinds are summed as
// angles with one sincos.  Each table entry is the cosine and sine of the entry's angle sum (host
// FP64), so the device does a few loads and complex products in place of a sum per term and a
// sincos.  The factor goes to `pre` (before the amplitudes are waited for), its use to `o`.
void emit_phase_tables(std::ostringstream& o, std::ostringstream& pre, const std::string& v, uint32_t key,
                       const TileGate& g, const TileProgram& tp, bool c64, std::vector<double>& coefs,
                       const std::string& flip)
{
    const bool uniform = key >= 20;
    struct PT
    {
        uint32_t lane_ctrl, lane_sel;
        uint64_t glb_ctrl, glb_sel;
        double th1, th0;
    };
    std::vector<PT> lane, glb, mixed;
    for (uint64_t j = 0; j < g.glb_sel; ++j) {
        const TileTerm& t = tp.terms[g.lane_sel + j];
        const auto d = term_values(t, c64);
        PT x {t.lane_ctrl, uniform ? t.lane_sel : 0u, t.glb_ctrl, uniform && t.lane_sel == 0 ? t.glb_sel : 0ull,
              std::arg(d[1]), d[0] == std::complex<double> {1.0, 0.0} ? 0.0 : std::arg(d[0])};
        if (!uniform) {
            x.th0 = 0.0;   // a run on a register bit: the factor multiplies the slots whose bit is 1
        }
        const bool lane_only = x.glb_ctrl == 0 && x.glb_sel == 0;
        const bool glb_only = x.lane_ctrl == 0 && x.lane_sel == 0;
        if (lane_only) {
            lane.push_back(x);
        }
        else if (glb_only) {
            glb.push_back(x);
        }
        else {
            mixed.push_back(x);
        }
    }
    // angle of a term at given thread / outside bits
    const auto angle = [&](const PT& x, uint32_t tid, uint64_t base) {
        const bool ok = (tid & x.lane_ctrl) == x.lane_ctrl && (base & x.glb_ctrl) == x.glb_ctrl;
        if (!ok) {
            return 0.0;
        }
        if (!uniform) {
            return x.th1;
        }
        const bool sel = x.lane_sel ? (tid & x.lane_sel) != 0 : (base & x.glb_sel) != 0;
        return sel ? x.th1 : x.th0;
    };
    const auto push_entry = [&](double a) {
        coefs.push_back(std::cos(a));
        coefs.push_back(std::sin(a));
    };
    pre << "  double " << v << "s1, " << v << "c1;\n  {\n    double fr = 1.0, fi = 0.0;\n";
    const auto mul = [&](const std::string& xr, const std::string& xi) {
        pre << "    { const double xr = " << xr << ", xi = " << xi << "; const double nr = fr * xr - fi * xi; fi = fr * xi + fi * xr; fr = nr; }\n";
    };
    if (!lane.empty()) {
        const uint32_t nt = 1u << tp.n_thr_bits;
        const size_t off = coefs.size();
        for (uint32_t tid = 0; tid < nt; ++tid) {
            double a = 0.0;
            for (const PT& x : lane) {
                a += angle(x, tid, 0);
            }
            push_entry(a);
        }
        pre << "    fr = cf[" << off << " + 2 * tid]; fi = cf[" << off + 1 << " + 2 * tid];\n";
    }
    // outside-bit terms: connected groups of the bits they read, packed into tables of <= 6 bits
    std::vector<std::pair<uint64_t, std::vector<PT>>> comps;
    for (const PT& x : glb) {
        const uint64_t m = x.glb_ctrl | x.glb_sel;
        std::pair<uint64_t, std::vector<PT>> merged {m, {x}};
        for (size_t i = 0; i < comps.size();) {
            if (comps[i].first & merged.first) {
                merged.first |= comps[i].first;
                merged.second.insert(merged.second.end(), comps[i].second.begin(), comps[i].second.end());
                comps.erase(comps.begin() + static_cast<std::ptrdiff_t>(i));
            }
            else {
                ++i;
            }
        }
        comps.push_back(std::move(merged));
    }
    std::vector<std::pair<uint64_t, std::vector<PT>>> groups;
    for (auto& c : comps) {
        if (std::popcount(c.first) > 6) {
            mixed.insert(mixed.end(), c.second.begin(), c.second.end());
            continue;
        }
        bool placed = false;
        for (auto& gr : groups) {
            if (std::popcount(gr.first | c.first) <= 6) {
                gr.first |= c.first;
                gr.second.insert(gr.second.end(), c.second.begin(), c.second.end());
                placed = true;
                break;
            }
        }
        if (!placed) {
            groups.push_back(std::move(c));
        }
    }
    for (const auto& gr : groups) {
        std::vector<int> bits;
        for (int b = 0; b < 64; ++b) {
            if ((gr.first >> b) & 1ull) {
                bits.push_back(b);
            }
        }
        const size_t off = coefs.size();
        for (uint32_t idx = 0; idx < (1u << bits.size()); ++idx) {
            uint64_t base = 0;
            for (size_t i = 0; i < bits.size(); ++i) {
                if ((idx >> i) & 1u) {
                    base |= uint64_t {1} << bits[i];
                }
            }
            double a = 0.0;
            for (const PT& x : gr.second) {
                a += angle(x, 0, base);
            }
            push_entry(a);
        }
        std::string ix;
        for (size_t i = 0; i < bits.size(); ++i) {
            ix += (i ? " | " : "") + std::string {"(((base >> "} + std::to_string(bits[i]) + ") & 1ull) << " + std::to_string(i) + ")";
        }
        pre << "    { const uint32_t gi = static_cast<uint32_t>(" << ix << ");\n  ";
        mul("cf[" + std::to_string(off) + " + 2 * gi]", "cf[" + std::to_string(off + 1) + " + 2 * gi]");
        pre << "    }\n";
    }
    if (!mixed.empty()) {
        pre << "    { double a = 0.0;\n";
        for (const PT& x : mixed) {
            coefs.push_back(x.th1);
            const size_t i1 = coefs.size() - 1;
            const std::string ok = cond_expr(x.glb_ctrl, x.lane_ctrl);
            if (uniform) {
                coefs.push_back(x.th0);
                const std::string sel = sel_expr(x.lane_sel, x.glb_sel);
                pre << "      a += (" << ok << ") ? (" << sel << " ? cf[" << i1 << "] : cf[" << i1 + 1 << "]) : 0.0;\n";
            }
            else {
                pre << "      a += (" << ok << ") ? cf[" << i1 << "] : 0.0;\n";
            }
        }
        pre << "      double s, c; sincos(a, &s, &c);\n  ";
        mul("c", "s");
        pre << "    }\n";
    }
    pre << "    " << v << "c1 = fr; " << v << "s1 = fi;\n  }\n";
    const std::string c1 = "static_cast<R>(" + v + "c1)", s1 = "static_cast<R>(" + v + "s1)";
    if (uniform) {
        o << "  scale<R, " << (key - 20) << ">(v, " << c1 << ", " << s1 << ");\n";
    }
    else {
        emit_diag_apply(o, key, "1", "0", c1, s1, true, flip);
    }
}

// A merged diagonal run whose terms are all pure phases (|d0| = |d1| = 1, as for CP, RZ, T, S and
// P) as a sum of FP64 angles selected by each term's condition and one sincos per factor, in place
// of a complex product per term.  Entries exactly 1 are skipped (structure, like the conditions).
// The conditions read only the thread index and the tile's outside bits, so the angle sums and the
// sincos go to `pre`, which the kernel runs before it waits for the tile's amplitudes (the arithmetic
// overlaps the loads); the sums run in four accumulators (short dependency chains).  `o` gets the
// multiplication of the amplitudes.
bool emit_phase_run(std::ostringstream& o, std::ostringstream& pre, int id, const TileGate& g, const TileProgram& tp,
                    bool c64, std::vector<double>& coefs, const std::string& flip)
{
    // one FP64 sincos costs about as much as the complex products of eight terms
    constexpr uint64_t MIN_TERMS = 8;
    if (g.glb_sel < MIN_TERMS) {
        return false;
    }
    const uint32_t key = g.code - TV_MERGE_BASE;
    const bool uniform = key >= 20;
    bool any0 = false;
    for (uint64_t j = 0; j < g.glb_sel; ++j) {
        const auto d = term_values(tp.terms[g.lane_sel + j], c64);
        if (std::abs(std::abs(d[0]) - 1.0) > 1.0e-13 || std::abs(std::abs(d[1]) - 1.0) > 1.0e-13) {
            return false;
        }
        any0 = any0 || d[0] != std::complex<double> {1.0, 0.0};
    }
    const bool need0 = any0 && !uniform;
    const std::string v = "ph" + std::to_string(id) + "_";
    if (!need0) {
        emit_phase_tables(o, pre, v, key, g, tp, c64, coefs, flip);
        return true;
    }
    pre << "  double " << v << "s1, " << v << "c1" << (need0 ? ", " + v + "s0, " + v + "c0" : std::string {}) << ";\n"
        << "  {\n    double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};\n";
    for (uint64_t j = 0; j < g.glb_sel; ++j) {
        const TileTerm& t = tp.terms[g.lane_sel + j];
        const auto d = term_values(t, c64);
        const std::string ok = cond_expr(t.glb_ctrl, t.lane_ctrl);
        const std::string acc = "[" + std::to_string(j % 4) + "]";
        const auto angle = [&](std::complex<double> z) {
            coefs.push_back(std::arg(z));
            return "cf[" + std::to_string(coefs.size() - 1) + "]";
        };
        if (uniform) {
            const std::string sel = sel_expr(t.lane_sel, t.glb_sel);
            const bool one0 = d[0] == std::complex<double> {1.0, 0.0};
            const std::string p1 = angle(d[1]);
            const std::string p0 = one0 ? std::string {"0.0"} : angle(d[0]);
            pre << "    a1" << acc << " += (" << ok << ") ? (" << sel << " ? " << p1 << " : " << p0 << ") : 0.0;\n";
        }
        else {
            pre << "    if (" << ok << ") {";
            if (need0 && d[0] != std::complex<double> {1.0, 0.0}) {
                pre << " a0" << acc << " += " << angle(d[0]) << ";";
            }
            pre << " a1" << acc << " += " << angle(d[1]) << "; }\n";
        }
    }
    pre << "    sincos((a1[0] + a1[1]) + (a1[2] + a1[3]), &" << v << "s1, &" << v << "c1);\n";
    if (need0) {
        pre << "    sincos((a0[0] + a0[1]) + (a0[2] + a0[3]), &" << v << "s0, &" << v << "c0);\n";
    }
    pre << "  }\n";
    const std::string c1 = "static_cast<R>(" + v + "c1)", s1 = "static_cast<R>(" + v + "s1)";
    if (uniform) {
        o << "  scale<R, " << (key - 20) << ">(v, " << c1 << ", " << s1 << ");\n";
        return true;
    }
    if (need0) {
        emit_diag_apply(o, key, "static_cast<R>(" + v + "c0)", "static_cast<R>(" + v + "s0)", c1, s1, false, flip);
    }
    else {
        emit_diag_apply(o, key, "1", "0", c1, s1, true, flip);
    }
    return true;
}

}  // namespace

JitStats& jit_stats() { return cache().stats; }

void set_jit_mode(int mode)
{
    if (mode < 0 || mode > 2) {
        throw InvalidArg {"jit mode must be 0 (off), 1 (large states) or 2 (every tile pass)"};
    }
    std::lock_guard<std::mutex> lk(cache().mu);
    cache().mode = mode;
}

int jit_mode() { return cache().mode; }

void set_jit_tuning(const JitTuning& t)
{
    if (t.min_waves_per_simd < 0 || t.min_waves_per_simd > 8 || t.blocks_per_cu < 0 || t.blocks_per_cu > 64
        || t.priority < 0 || t.priority > 2 || t.nontemporal < 0 || t.nontemporal > 3) {
        throw InvalidArg {"jit tuning: min waves per SIMD must be in [0, 8], blocks per CU in [0, 64], priority "
                          "in [0, 2] and nontemporal mode in [0, 3]"};
    }
    std::lock_guard<std::mutex> lk(cache().mu);
    cache().tuning = t;
}

JitTuning jit_tuning() { return cache().tuning; }

bool jit_wanted(const kq_state* s)
{
    const int mode = cache().mode;
    return mode == 2 || (mode == 1 && s->nvec >= 18);
}

// Generate the kernel source of one forward tile pass and the layout of its coefficient buffer.
std::string jit_source(const TileProgram& tp, int precision, std::vector<double>& coefs, PassInput input)
{
    const bool from_basis = input == PassInput::basis;
    const bool skip_zero = input == PassInput::skip_zero;
    coefs.clear();
    const bool c64 = precision == KQ_C64;
    const int nt = 1 << tp.n_thr_bits;
    std::ostringstream o;
    const JitTuning tune = jit_tuning();
    o << "typedef unsigned long long uint64_t;\ntypedef unsigned int uint32_t;\ntypedef int int32_t;\n"
      << DEVICE_HEADER << "\nusing R = " << (c64 ? "float" : "double") << ";\n"
      << "extern \"C\" __global__ __launch_bounds__(" << nt << ")";
    if (tune.min_waves_per_simd > 0) {
        o << " __attribute__((amdgpu_waves_per_eu(" << tune.min_waves_per_simd << ")))";
    }
    // an out-of-place pass (the state's last pass putting every qubit on its final bit) loads from `a`
    // and stores into `dst`, at the destination bits of every tile and outside bit
    const bool oop = !tp.out_freepos.empty();
    // from_basis: the state is the basis vector |basis> not yet written; the pass builds its tiles from
    // that index instead of loading them
    // skip_zero: the workgroups whose outside bits under zmask differ from zval hold a zero tile
    o << " void kq_pass(kq::dev::Amp<R>* __restrict__ a, const double* __restrict__ cf, uint32_t n_blocks"
      << (oop ? ", kq::dev::Amp<R>* __restrict__ dst" : "") << (from_basis ? ", uint64_t basis" : "")
      << (skip_zero ? ", uint64_t zmask, uint64_t zval" : "") << ")\n{\n"
      << "  using namespace kq::dev;\n";
    bool remaps = tp.out_layout != 0;
    for (const TileStep& st : tp.steps) {
        remaps = remaps || st.layout != 0;
    }
    if (remaps) {
        o << "  __shared__ " << (c64 ? "Amp<R>" : "R") << " tile[" << (1 << tp.k) << "];\n"
          << "  char* const lds = reinterpret_cast<char*>(tile);\n";
    }
    // Blocks walk the tiles with a grid stride.  With more tiles than blocks, the next tile's amplitudes
    // are loaded into registers before the current tile's gates run, so HBM reads overlap the arithmetic.
    const bool persistent = tune.blocks_per_cu > 0;
    if (persistent && input != PassInput::load) {
        throw InvalidArg {"jit: a pass from a basis state or over zero tiles has no persistent form"};
    }
    o << "  const uint32_t tid = threadIdx.x;\n  Cx<R> v[16];\n";
    if (persistent) {
        o << "  const uint64_t thr_in = " << hbm_thread_expr(tp.layouts[0], tp.n_thr_bits) << ";\n"
          << "  uint32_t b = blockIdx.x;\n  Amp<R> nx[16];\n"
          << "  const auto fetch = [&](uint64_t blk) {\n    const uint64_t off = (" << deposit_expr(tp.freepos)
          << ") | thr_in;\n";
        for (int r = 0; r < 16; ++r) {
            o << "    nx[" << r << "] = a[off + " << hex64(tp.layouts[0].hbm_slot[r]) << "];\n";
        }
        o << "  };\n  fetch(b);\n  for (; b < n_blocks; b += gridDim.x) {\n"
          << "    const uint64_t blk = b;\n    const uint64_t base = " << deposit_expr(tp.freepos) << ";\n";
        for (int r = 0; r < 16; ++r) {
            o << "    v[" << r << "] = {nx[" << r << "].x, nx[" << r << "].y};\n";
        }
        o << "    if (b + gridDim.x < n_blocks) {\n      fetch(b + gridDim.x);\n    }\n";
    }
    else {
        // Workgroups are dealt to the 8 XCDs round robin; workgroup i takes tile (i >> 3) | (i & 7) << (NB - 3),
        // so each XCD walks a contiguous eighth of the tiles and the outside bits that vary fastest on
        // one XCD are the low ones (every HBM channel busy), not the ones above the XCD index.
        const size_t nb = tp.freepos.size();
        o << "  {\n    const uint64_t blk = ";
        if (nb >= 3) {
            o << "(uint64_t {blockIdx.x} >> 3) | (uint64_t {blockIdx.x & 7u} << " << nb - 3 << ")";
        }
        else {
            o << "blockIdx.x";
        }
        o << ";\n    const uint64_t base = " << deposit_expr(tp.freepos) << ";\n";
    }
    // Controlled X gates whose control is a thread or block bit (a register swap under a lane- or
    // block-dependent condition) are not executed in registers: each stays pending until the next
    // LDS or HBM write, where it becomes an XOR of the write address, unless a gate on its register
    // bit comes first.  Register slot r of layout L sits at slot address L.*_slot[r], which is
    // XOR-linear in r, so swapping slots r and r ^ (1 << P) is the address XOR L.*_slot[1 << P].
    struct PendingX
    {
        int P;
        std::string cond;
        uint64_t glb_ctrl;
        uint32_t lane_ctrl;
    };
    std::vector<PendingX> pending;
    const auto linear = [](const auto& slots) {
        for (int r = 0; r < 16; ++r) {
            for (int q = 0; q < 16; ++q) {
                if (slots[r ^ q] != (slots[r] ^ slots[q])) {
                    return false;
                }
            }
        }
        return true;
    };
    const auto pending_xor = [&](const auto& slots, const char* zero) {
        std::string e = zero;
        for (const PendingX& x : pending) {
            e += " ^ ((" + x.cond + ") ? " + (std::string {zero} == "0ull" ? hex64(slots[1 << x.P]) : hex32(static_cast<uint32_t>(slots[1 << x.P])))
               + " : " + zero + ")";
        }
        return e;
    };
    const auto io = [&](const TileLayout& L, bool load) {
        const bool to_dst = oop && !load;
        o << "  {\n    const uint64_t off = (" << (to_dst ? "(" + deposit_expr(tp.out_freepos) + ")" : std::string {"base"})
          << " | (" << hbm_thread_expr(L, tp.n_thr_bits) << "))";
        if (!load) {
            o << " ^ (" << pending_xor(L.hbm_slot, "0ull") << ")";
            pending.clear();
        }
        o << ";\n";
        const int ntm = nontemporal_mode(tune.nontemporal, static_cast<uint32_t>(tp.k + tp.freepos.size()), c64 ? 8 : 16);
        const char* nt_ld = ntm > 0 ? "true" : "false";
        const char* nt_st = ntm > 1 ? "true" : "false";
        for (int r = 0; r < 16; ++r) {
            if (load) {
                o << "    { const Amp<R> x = amp_load<" << nt_ld << ">(a + (off ^ " << hex64(L.hbm_slot[r]) << ")); v[" << r
                  << "] = {x.x, x.y}; }\n";
            }
            else {
                o << "    { Amp<R> x; x.x = v[" << r << "].re; x.y = v[" << r << "].im; amp_store<" << nt_st << ">("
                  << (to_dst ? "dst" : "a") << " + (off ^ " << hex64(L.hbm_slot[r]) << "), x); }\n";
            }
        }
        o << "  }\n";
    };
    const auto remap = [&](const TileLayout& A0, const TileLayout& B0) {
        std::vector<std::pair<int, int>> lane_pending;
        for (const PendingX& x : pending) {
            if (x.glb_ctrl == 0 && std::popcount(x.lane_ctrl) == 1) {
                lane_pending.emplace_back(std::countr_zero(x.lane_ctrl), x.P);
            }
        }
        // complex128 remaps move the real parts, then the imaginary parts, through a buffer of 8-byte
        // entries (half the LDS of a whole tile: more workgroups per CU); complex64 amplitudes are
        // 8-byte entries already
        const auto [A, B] = bank_spread(A0, B0, tp.k, tp.n_thr_bits, amp_bytes(precision), 8, BANKS_8B, lane_pending);
        o << "  {\n    const uint32_t xa = (" << lds_thread_expr(A, tp.n_thr_bits) << ") ^ ("
          << pending_xor(A.lds_slot, "0u") << ");\n"
          << "    const uint32_t xb = " << lds_thread_expr(B, tp.n_thr_bits) << ";\n";
        pending.clear();
        if (c64) {
            o << "    lds_barrier();\n";
            for (int r = 0; r < 16; ++r) {
                o << "    { Amp<R> x; x.x = v[" << r << "].re; x.y = v[" << r << "].im; *reinterpret_cast<Amp<R>*>(lds + (xa ^ "
                  << hex32(A.lds_slot[r]) << ")) = x; }\n";
            }
            o << "    lds_barrier();\n";
            for (int r = 0; r < 16; ++r) {
                o << "    { const Amp<R> x = *reinterpret_cast<const Amp<R>*>(lds + (xb ^ " << hex32(B.lds_slot[r])
                  << ")); v[" << r << "] = {x.x, x.y}; }\n";
            }
        }
        else {
            for (const char* part : {"re", "im"}) {
                o << "    lds_barrier();\n";
                for (int r = 0; r < 16; ++r) {
                    o << "    *reinterpret_cast<R*>(lds + (xa ^ " << hex32(A.lds_slot[r]) << ")) = v[" << r << "]." << part
                      << ";\n";
                }
                o << "    lds_barrier();\n";
                for (int r = 0; r < 16; ++r) {
                    o << "    v[" << r << "]." << part << " = *reinterpret_cast<const R*>(lds + (xb ^ " << hex32(B.lds_slot[r])
                      << "));\n";
                }
            }
        }
        o << "  }\n";
    };
    const auto push_gate_coefs = [&](const TileGate& g) {
        const size_t off = coefs.size();
        if (c64) {
            for (int i = 0; i < 4; ++i) {
                coefs.push_back(g.m[i]);   // 8 packed floats
            }
        }
        else {
            coefs.insert(coefs.end(), g.m, g.m + 8);
        }
        return off;
    };
    const auto push_term_coefs = [&](const TileTerm& t) {
        const size_t off = coefs.size();
        if (c64) {
            coefs.push_back(t.d[0]);
            coefs.push_back(t.d[1]);
        }
        else {
            coefs.insert(coefs.end(), t.d, t.d + 4);
        }
        return off;
    };
    // zeros stored where the pass stores, by a workgroup whose tile is 0 (and stays 0 under the pass's
    // gates, which are linear and act within each tile), which then ends
    const auto zero_store = [&](const std::string& cond) {
        const TileLayout& S = tp.layouts[tp.out_layout];
        const int ntm = nontemporal_mode(tune.nontemporal, static_cast<uint32_t>(tp.k + tp.freepos.size()), c64 ? 8 : 16);
        o << "  if (" << cond << ") {\n"
          << "    const uint64_t zoff = " << (oop ? "(" + deposit_expr(tp.out_freepos) + ")" : std::string {"base"}) << " | ("
          << hbm_thread_expr(S, tp.n_thr_bits) << ");\n"
          << "    Amp<R> zero;\n    zero.x = R(0);\n    zero.y = R(0);\n";
        for (int r = 0; r < 16; ++r) {
            o << "    amp_store<" << (ntm > 1 ? "true" : "false") << ">(" << (oop ? "dst" : "a") << " + (zoff ^ "
              << hex64(S.hbm_slot[r]) << "), zero);\n";
        }
        o << "    return;\n  }\n";
    };
    if (skip_zero) {
        if (oop) {
            zero_store("(base & zmask) != zval");   // the destination buffer holds no zeros yet
        }
        else {
            // in place: the buffer already holds the zeros
            o << "  if ((base & zmask) != zval) {\n    return;\n  }\n";
        }
    }
    if (from_basis) {
        // Every tile but the one holding the basis amplitude is 0: those workgroups store zeros.  The one
        // that holds it builds its tile, amplitude (off ^ slot) 1 at the basis index and 0 elsewhere, and
        // runs the pass.  Nothing is read.
        uint64_t free_mask = 0;
        for (const uint8_t q : tp.freepos) {
            free_mask |= uint64_t {1} << q;
        }
        zero_store("(basis & " + hex64(free_mask) + ") != base");
        const TileLayout& L = tp.layouts[0];
        o << "  {\n    const uint64_t off = base | (" << hbm_thread_expr(L, tp.n_thr_bits) << ");\n";
        for (int r = 0; r < 16; ++r) {
            o << "    v[" << r << "] = {((off ^ " << hex64(L.hbm_slot[r]) << ") == basis) ? R(1) : R(0), R(0)};\n";
        }
        o << "  }\n";
    }
    else if (!persistent) {
        if (tune.priority > 0) {
            o << "  wave_priority<3>();\n";
        }
        io(tp.layouts[0], true);
        if (tune.priority > 0) {
            o << "  wave_priority<0>();\n";
        }
    }
    // pivots divided out of normalised gates, multiplied back in once before the store
    const bool normalise = tune.normalise;
    std::complex<double> factor {1.0, 0.0};
    bool factor_complex = false;
    bool normalised = false;
    bool all_linear = true;
    for (const TileLayout& L : tp.layouts) {
        all_linear = all_linear && linear(L.lds_slot) && linear(L.hbm_slot);
    }
    // register bits a record reads or writes (target and register control)
    const auto touched = [](const TileGate& g) {
        uint32_t bits = 0;
        uint32_t P = 4, C = 0;
        if (g.code >= TV_MERGE_BASE) {
            const uint32_t key = g.code - TV_MERGE_BASE;
            if (key >= 20) {
                C = key - 20;
            }
            else {
                P = key / 5;
                C = key % 5;
            }
        }
        else if (g.code >= TV_UNIFORM_BASE) {
            C = g.code - TV_UNIFORM_BASE;
        }
        else {
            P = (g.code / 5) % 4;
            C = g.code % 5;
        }
        if (P < 4) {
            bits |= 1u << P;
        }
        if (C > 0) {
            bits |= 1u << (C - 1);
        }
        return bits;
    };
    const auto materialise = [&](uint32_t bits) {
        for (size_t i = 0; i < pending.size();) {
            if ((bits >> pending[i].P) & 1u) {
                o << "  if (" << pending[i].cond << ") {\n    const Coef<R> m {};\n    apply<R, K_SWAP, " << pending[i].P
                  << ", 0>(v, m);\n  }\n";
                pending.erase(pending.begin() + static_cast<std::ptrdiff_t>(i));
            }
            else {
                ++i;
            }
        }
    };
    // what precedes the gates (declarations, the loads); phase sums emitted into `pre` go after it
    const std::string head = o.str();
    o.str("");
    o.clear();
    std::ostringstream pre;
    int n_runs = 0;
    uint32_t cur = 0;
    for (const TileStep& st : tp.steps) {
        if (st.layout != cur) {
            remap(tp.layouts[cur], tp.layouts[st.layout]);
            cur = st.layout;
        }
        for (uint32_t gi = st.g_begin; gi < st.g_begin + st.g_count; ++gi) {
            const TileGate& g = tp.gates[gi];
            uint32_t bits = touched(g);
            if (g.code < TV_UNIFORM_BASE) {
                const uint32_t kind = g.code / 20;
                const int Pg = static_cast<int>((g.code / 5) % 4);
                const uint32_t pbit = 1u << Pg;
                std::string flip;
                for (const PendingX& x : pending) {
                    if (x.P == Pg) {
                        flip += (flip.empty() ? "(" : " != (") + x.cond + ")";
                    }
                }
                if (kind == TK_SWAP) {
                    // an X on the same register bit (whatever controls it) commutes with the pending ones
                    bits &= ~pbit;
                }
                else if (!flip.empty() && g.code % 5 == 0 && g.lane_ctrl == 0 && g.glb_ctrl == 0) {
                    // the pending X gates stay pending: where they apply, the lanes run X G X instead of G
                    materialise(bits & ~pbit);
                    if (normalise && (kind == TK_REAL || kind == TK_AXIS || kind == TK_DENSE)
                        && emit_normalised(o, g, c64, coefs, factor, factor_complex, flip)) {
                        normalised = true;
                        continue;
                    }
                    const size_t off = push_gate_coefs(g);
                    o << "  { const Coef<R> m = flip_coef<R>(coef_at<R>(cf + " << off << "), " << flip << "); apply<R, "
                      << (kind == TK_DIAG1 ? static_cast<uint32_t>(TK_DIAG) : kind) << ", " << Pg << ", 0>(v, m); }\n";
                    continue;
                }
            }
            // a merged diagonal run on a register bit with controlled X gates pending on it runs flipped
            std::string run_flip;
            if (g.code >= TV_MERGE_BASE && g.code - TV_MERGE_BASE < 20 && (g.code - TV_MERGE_BASE) % 5 == 0) {
                const int Pm = static_cast<int>((g.code - TV_MERGE_BASE) / 5);
                for (const PendingX& x : pending) {
                    if (x.P == Pm) {
                        run_flip += (run_flip.empty() ? "(" : " != (") + x.cond + ")";
                    }
                }
                if (!run_flip.empty()) {
                    bits &= ~(1u << Pm);
                }
            }
            materialise(bits);
            if (all_linear && g.code < TV_UNIFORM_BASE && g.code / 20 == TK_SWAP && g.code % 5 == 0
                && (g.lane_ctrl != 0 || g.glb_ctrl != 0)) {
                pending.push_back(PendingX {static_cast<int>((g.code / 5) % 4), cond_expr(g.glb_ctrl, g.lane_ctrl), g.glb_ctrl, g.lane_ctrl});
                continue;
            }
            if (g.code >= TV_MERGE_BASE && emit_phase_run(o, pre, n_runs, g, tp, c64, coefs, run_flip)) {
                ++n_runs;
                continue;
            }
            if (g.code >= TV_MERGE_BASE) {
                const uint32_t key = g.code - TV_MERGE_BASE;
                const bool uniform = key >= 20;
                o << "  {\n    R f0r = 1, f0i = 0, f1r = 1, f1i = 0;\n";
                for (uint64_t j = 0; j < g.glb_sel; ++j) {
                    const TileTerm& t = tp.terms[g.lane_sel + j];
                    const size_t off = push_term_coefs(t);
                    o << "    { R d0r, d0i, d1r, d1i; term_at<R>(cf + " << off << ", d0r, d0i, d1r, d1i); const bool ok = "
                      << cond_expr(t.glb_ctrl, t.lane_ctrl) << ";";
                    if (uniform) {
                        o << " const bool sel = " << sel_expr(t.lane_sel, t.glb_sel)
                          << "; mul_if<R>(ok, f1r, f1i, sel ? d1r : d0r, sel ? d1i : d0i); }\n";
                    }
                    else {
                        o << " mul_if<R>(ok, f0r, f0i, d0r, d0i); mul_if<R>(ok, f1r, f1i, d1r, d1i); }\n";
                    }
                }
                if (uniform) {
                    o << "    scale<R, " << (key - 20) << ">(v, f1r, f1i);\n  }\n";
                }
                else {
                    emit_diag_apply(o, key, "f0r", "f0i", "f1r", "f1i", false, run_flip);
                    o << "  }\n";
                }
                continue;
            }
            if (normalise && emit_normalised(o, g, c64, coefs, factor, factor_complex)) {
                normalised = true;
                continue;
            }
            const size_t off = push_gate_coefs(g);
            o << "  if (" << cond_expr(g.glb_ctrl, g.lane_ctrl) << ") {\n    const Coef<R> m = coef_at<R>(cf + " << off
              << ");\n";
            if (g.code >= TV_UNIFORM_BASE) {
                o << "    const bool sel = " << sel_expr(g.lane_sel, g.glb_sel) << ";\n    scale<R, "
                  << (g.code - TV_UNIFORM_BASE) << ">(v, sel ? m.m[6] : m.m[0], sel ? m.m[7] : m.m[1]);\n  }\n";
            }
            else {
                o << "    apply<R, " << g.code / 20 << ", " << (g.code / 5) % 4 << ", " << g.code % 5 << ">(v, m);\n  }\n";
            }
        }
    }
    if (normalised) {
        if (factor_complex) {
            coefs.push_back(factor.real());
            coefs.push_back(factor.imag());
            o << "  scale<R, 0>(v, static_cast<R>(cf[" << coefs.size() - 2 << "]), static_cast<R>(cf[" << coefs.size() - 1
              << "]));\n";
        }
        else {
            coefs.push_back(factor.real());
            o << "  scale_real<R>(v, static_cast<R>(cf[" << coefs.size() - 1 << "]));\n";
        }
    }
    if (tp.out_layout != cur) {
        const TileLayout& A = tp.layouts[cur];
        const TileLayout& B = tp.layouts[tp.out_layout];
        const bool same_lds = std::equal(A.lds_slot, A.lds_slot + 16, B.lds_slot)
                           && std::equal(A.lds_thr, A.lds_thr + 8, B.lds_thr);
        if (!same_lds) {
            remap(A, B);
        }
        else if (!oop) {
            // Same registers, other HBM offsets: the store moves qubits to other bits of the tile, so
            // a thread writes amplitudes other threads of the workgroup read.  Every load of the
            // workgroup must have completed first: a full barrier (it waits for outstanding loads).
            o << "  __syncthreads();\n";
        }
        cur = tp.out_layout;
    }
    if (tune.priority > 1) {
        o << "  wave_priority<2>();\n";
    }
    io(tp.layouts[cur], false);
    o << "  }\n}\n";
    return head + pre.str() + o.str();
}

// The coefficient buffer of an adjoint reverse pass in the order its generated kernel reads it: per
// record (steps in order, each step's records in order) the terms of a merged diagonal run, or the 8
// values of a uniform diagonal, or the normalised values of an uncontrolled one-qubit gate (its pivot
// divided out, as in the forward passes, the pivot going into the pass factor F), or the 8 values of any
// other gate, then the 8 values of its derivative record if it has one; finally F itself, which
// multiplies both vectors before the store.  complex64 values as packed floats.  offs (optional)
// receives the offsets of the terms, uniform and general gate values and derivative values in that
// order (normalised gates carry their offsets in their statements).  Depends on the coefficients only,
// so a pass of known structure and forms needs no new source.
void adjoint_coefs(const TileProgram& tp, int precision, AdjointCoefs& out, std::vector<size_t>* offs)
{
    const bool c64 = precision == KQ_C64;
    std::vector<double>& coefs = out.coefs;
    coefs.clear();
    out.dscale.assign(tp.derivs.size(), 1.0);
    out.forms.assign(tp.gates.size(), std::string {});
    if (offs) {
        offs->clear();
    }
    const auto push8 = [&](const double* m) {
        if (offs) {
            offs->push_back(coefs.size());
        }
        if (c64) {
            for (int i = 0; i < 4; ++i) {
                coefs.push_back(m[i]);   // 8 packed floats
            }
        }
        else {
            coefs.insert(coefs.end(), m, m + 8);
        }
    };
    // Products of psi and lambda grow by 1 / |F|^2 inside the pass: complex64 keeps |F| >= 1e-12, so they
    // stay far below the FP32 range.
    const double floor = c64 ? 1.0e-12 : 1.0e-200;
    std::complex<double> factor {1.0, 0.0};
    bool factor_complex = false;
    bool normalised = false;
    for (const TileStep& st : tp.steps) {
        for (uint32_t gi = st.g_begin; gi < st.g_begin + st.g_count; ++gi) {
            const TileGate& g = tp.gates[gi];
            if (g.pad != 0) {
                out.dscale[g.pad - 1] = std::norm(factor);   // |F|^2 where the derivative is taken
            }
            if (g.code >= TV_MERGE_BASE) {
                for (uint64_t j = 0; j < g.glb_sel; ++j) {
                    const TileTerm& t = tp.terms[g.lane_sel + j];
                    if (offs) {
                        offs->push_back(coefs.size());
                    }
                    if (c64) {
                        coefs.push_back(t.d[0]);
                        coefs.push_back(t.d[1]);
                    }
                    else {
                        coefs.insert(coefs.end(), t.d, t.d + 4);
                    }
                }
            }
            else if (g.code >= TV_UNIFORM_BASE) {
                double m[8];
                const auto mv = record_values(g, c64);
                std::copy(mv.begin(), mv.end(), m);
                push8(c64 ? reinterpret_cast<const double*>(g.m) : m);
            }
            else {
                std::ostringstream form;
                if (emit_normalised(form, g, c64, coefs, factor, factor_complex, {}, {"v", "w"}, floor)) {
                    out.forms[gi] = form.str();
                    normalised = true;
                }
                else {
                    push8(g.m);
                }
            }
            if (g.pad != 0) {
                push8(tp.derivs[g.pad - 1].m);
            }
        }
    }
    if (normalised) {
        coefs.push_back(factor.real());
        coefs.push_back(factor.imag());
    }
}

// Structure of an adjoint reverse pass: everything its generated source depends on besides the
// coefficients (records' variants, conditions and selections, derivative generators, steps, layouts,
// outside bits), with the precision and the kernel tuning; passes with equal keys have equal sources.
std::vector<uint64_t> adjoint_structure_key(const TileProgram& tp, int precision, const AdjointCoefs& c)
{
    const JitTuning t = jit_tuning();
    std::vector<uint64_t> key {static_cast<uint64_t>(precision), static_cast<uint64_t>(tp.k), static_cast<uint64_t>(tp.n_thr_bits),
                               tp.out_layout, static_cast<uint64_t>(t.nontemporal), static_cast<uint64_t>(t.priority),
                               static_cast<uint64_t>(t.min_waves_per_simd), tp.gates.size(), tp.steps.size(), tp.layouts.size()};
    const auto rec = [&](const TileGate& g) {
        key.push_back((uint64_t {g.code} << 32) | g.lane_ctrl);
        key.push_back((uint64_t {g.lane_sel} << 32) | g.pad);
        key.push_back(g.glb_ctrl);
        key.push_back(g.glb_sel);
    };
    for (const TileGate& g : tp.gates) {
        rec(g);
    }
    for (const TileGate& d : tp.derivs) {
        rec(d);
        for (const double v : record_values(d, precision == KQ_C64)) {
            key.push_back(std::bit_cast<uint64_t>(v));   // the generator's form is chosen from these
        }
    }
    for (const TileTerm& m : tp.terms) {
        key.push_back((uint64_t {m.lane_ctrl} << 32) | m.lane_sel);
        key.push_back(m.glb_ctrl);
        key.push_back(m.glb_sel);
    }
    for (const TileStep& st : tp.steps) {
        key.push_back((uint64_t {st.g_begin} << 32) | st.g_count);
        key.push_back(st.layout);
    }
    for (const TileLayout& L : tp.layouts) {
        const auto* w = reinterpret_cast<const uint32_t*>(&L);
        for (size_t i = 0; i < sizeof(TileLayout) / sizeof(uint32_t); ++i) {
            key.push_back(w[i]);
        }
    }
    for (const uint8_t b : tp.freepos) {
        key.push_back(b);
    }
    // the normalised gates' statements (their form, and where their values sit in the buffer)
    for (const std::string& f : c.forms) {
        key.push_back(f.size());
        for (size_t i = 0; i < f.size(); i += 8) {
            uint64_t w = 0;
            std::memcpy(&w, f.data() + i, std::min<size_t>(8, f.size() - i));
            key.push_back(w);
        }
    }
    return key;
}

// Generate the kernel source of one adjoint reverse pass (two vectors: psi `a`, lambda `b`) and the
// layout of its coefficient buffer.  The same records as the interpreted adjoint tile kernel, as
// straight-line code: for every record, if it has a derivative record Q, Re <lambda| Q |psi> over the
// tile, reduced per wavefront into partials[slot * n_part + column]; then psi <- G psi, lambda <- G lambda.
// Remaps move psi, then lambda, through one LDS tile of whole amplitudes.  Uncontrolled one-qubit gates
// run with their pivot divided out (adjoint_coefs), the product of the pivots multiplying both vectors
// before the store.
std::string jit_adjoint_source(const TileProgram& tp, int precision, AdjointCoefs& ac)
{
    const JitTuning tune = jit_tuning();
    const bool c64 = precision == KQ_C64;
    const int nt = 1 << tp.n_thr_bits;
    const int waves = (nt + 63) / 64;
    const int width = nt < 64 ? nt : 64;
    std::ostringstream o;
    o << "typedef unsigned long long uint64_t;\ntypedef unsigned int uint32_t;\ntypedef int int32_t;\n"
      << DEVICE_HEADER << "\nusing R = " << (c64 ? "float" : "double") << ";\n"
      << "__device__ __forceinline__ double wsum(double x)\n{\n"
      << "  for (int off = 1; off < " << width << "; off <<= 1) {\n    x += __shfl_xor(x, off, 64);\n  }\n  return x;\n}\n"
      << "extern \"C\" __global__ __launch_bounds__(" << nt << ") void kq_adjoint_pass(kq::dev::Amp<R>* __restrict__ a, "
      << "kq::dev::Amp<R>* __restrict__ b, const double* __restrict__ cf, double* __restrict__ partials, uint32_t n_part)\n{\n"
      << "  using namespace kq::dev;\n";
    bool remaps = tp.out_layout != 0;
    for (const TileStep& st : tp.steps) {
        remaps = remaps || st.layout != 0;
    }
    if (remaps) {
        o << "  __shared__ Amp<R> tile[" << (1 << tp.k) << "];\n  char* const lds = reinterpret_cast<char*>(tile);\n";
    }
    const size_t nb = tp.freepos.size();
    o << "  const uint32_t tid = threadIdx.x;\n  Cx<R> v[16];\n  Cx<R> w[16];\n"
      << "  const uint64_t blk = ";
    if (nb >= 3) {
        o << "(uint64_t {blockIdx.x} >> 3) | (uint64_t {blockIdx.x & 7u} << " << nb - 3 << ")";
    }
    else {
        o << "blockIdx.x";
    }
    o << ";\n  const uint64_t base = " << deposit_expr(tp.freepos) << ";\n"
      << "  const uint64_t col = uint64_t {blockIdx.x} * " << waves << "u + (tid >> 6);\n";
    // Derivative partials, complex64 passes of whole wavefronts: each lane writes its value into a
    // wave-private LDS area (16 derivatives x 64 lanes); every 16 derivatives and at the end the wave
    // sums them together, 4 lanes per derivative over 16 values each, then 2 shuffle steps, instead of a
    // 6-step shuffle tree per derivative.  The fences order the area's writes and reads within the wave.
    const bool staged = c64 && nt >= 64;
    constexpr int STAGE = 16;
    std::vector<uint32_t> pend;
    if (staged) {
        o << "  __shared__ double dpart[" << waves * STAGE * 64 << "];\n"
          << "  double* const dp = dpart + (tid >> 6) * " << STAGE * 64 << "u;\n  const uint32_t lane = tid & 63u;\n";
    }
    const auto flush = [&] {
        if (pend.empty()) {
            return;
        }
        o << "  {\n    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"workgroup\");\n"
          << "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"workgroup\");\n"
          << "    const uint32_t j = lane >> 2;\n    double s = 0.0;\n"
          << "    if (j < " << pend.size() << "u) {\n      const double* q = dp + j * 64u + (lane & 3u) * 16u;\n";
        for (int i = 0; i < 16; ++i) {
            o << "      s += q[" << i << "];\n";
        }
        o << "    }\n    s += __shfl_xor(s, 1, 64);\n    s += __shfl_xor(s, 2, 64);\n"
          << "    if ((lane & 3u) == 0u && j < " << pend.size() << "u) {\n      uint32_t slot = " << pend[0] << "u;\n";
        for (size_t i = 1; i < pend.size(); ++i) {
            o << "      slot = j == " << i << "u ? " << pend[i] << "u : slot;\n";
        }
        o << "      partials[uint64_t {slot} * n_part + col] = s;\n    }\n"
          << "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"workgroup\");\n  }\n";
        pend.clear();
    };
    const auto io = [&](const TileLayout& L, bool load) {
        o << "  {\n    const uint64_t off = base | (" << hbm_thread_expr(L, tp.n_thr_bits) << ");\n";
        const int ntm = nontemporal_mode(tune.nontemporal, static_cast<uint32_t>(tp.k + tp.freepos.size()), c64 ? 8 : 16);
        const char* nt_ld = ntm > 0 ? "true" : "false";
        const char* nt_st = ntm > 1 ? "true" : "false";
        for (int r = 0; r < 16; ++r) {
            const std::string at = "(off ^ " + hex64(L.hbm_slot[r]) + ")";
            if (load) {
                o << "    { const Amp<R> x = amp_load<" << nt_ld << ">(a + " << at << "); v[" << r << "] = {x.x, x.y}; }\n"
                  << "    { const Amp<R> y = amp_load<" << nt_ld << ">(b + " << at << "); w[" << r << "] = {y.x, y.y}; }\n";
            }
            else {
                o << "    { Amp<R> x; x.x = v[" << r << "].re; x.y = v[" << r << "].im; amp_store<" << nt_st << ">(a + " << at << ", x); }\n"
                  << "    { Amp<R> y; y.x = w[" << r << "].re; y.y = w[" << r << "].im; amp_store<" << nt_st << ">(b + " << at << ", y); }\n";
            }
        }
        o << "  }\n";
    };
    const auto remap = [&](const TileLayout& A, const TileLayout& B) {
        o << "  {\n    const uint32_t xa = " << lds_thread_expr(A, tp.n_thr_bits) << ";\n    const uint32_t xb = "
          << lds_thread_expr(B, tp.n_thr_bits) << ";\n";
        for (const char* u : {"v", "w"}) {
            o << "    lds_barrier();\n";
            for (int r = 0; r < 16; ++r) {
                o << "    { Amp<R> x; x.x = " << u << "[" << r << "].re; x.y = " << u << "[" << r << "].im; "
                  << "*reinterpret_cast<Amp<R>*>(lds + (xa ^ " << hex32(A.lds_slot[r]) << ")) = x; }\n";
            }
            o << "    lds_barrier();\n";
            for (int r = 0; r < 16; ++r) {
                o << "    { const Amp<R> x = *reinterpret_cast<const Amp<R>*>(lds + (xb ^ " << hex32(B.lds_slot[r]) << ")); "
                  << u << "[" << r << "] = {x.x, x.y}; }\n";
            }
        }
        o << "  }\n";
    };
    // offsets of the coefficient buffer, consumed in the order adjoint_coefs laid it out
    std::vector<size_t> offs;
    adjoint_coefs(tp, precision, ac, &offs);
    size_t next_off = 0;
    const auto take = [&]() { return offs.at(next_off++); };
    // the statement applying record g to vector u (coefficients / factors already in scope)
    const auto apply_to = [&](const TileGate& g, const char* u) {
        std::ostringstream a;
        if (g.code >= TV_MERGE_BASE) {
            const uint32_t key = g.code - TV_MERGE_BASE;
            if (key >= 20) {
                a << "scale<R, " << (key - 20) << ">(" << u << ", f1r, f1i);";
            }
            else {
                a << "apply<R, K_DIAG, " << key / 5 << ", " << key % 5 << ">(" << u << ", fm);";
            }
        }
        else if (g.code >= TV_UNIFORM_BASE) {
            a << "scale<R, " << (g.code - TV_UNIFORM_BASE) << ">(" << u << ", fr, fi);";
        }
        else {
            a << "apply<R, " << g.code / 20 << ", " << (g.code / 5) % 4 << ", " << g.code % 5 << ">(" << u << ", m);";
        }
        return a.str();
    };
    if (tp.layouts.empty()) {
        throw InvalidArg {"jit adjoint: program without layouts"};
    }
    io(tp.layouts[0], true);
    uint32_t cur = 0;
    for (const TileStep& st : tp.steps) {
        if (st.layout != cur) {
            remap(tp.layouts[cur], tp.layouts[st.layout]);
            cur = st.layout;
        }
        for (uint32_t gi = st.g_begin; gi < st.g_begin + st.g_count; ++gi) {
            const TileGate& g = tp.gates[gi];
            o << "  {\n    const bool ok = " << cond_expr(g.glb_ctrl, g.lane_ctrl) << ";\n";
            if (g.code >= TV_MERGE_BASE) {
                const uint32_t key = g.code - TV_MERGE_BASE;
                const bool uniform = key >= 20;
                o << "    R f0r = 1, f0i = 0, f1r = 1, f1i = 0;\n";
                for (uint64_t j = 0; j < g.glb_sel; ++j) {
                    const TileTerm& t = tp.terms[g.lane_sel + j];
                    const size_t off = take();
                    o << "    { R d0r, d0i, d1r, d1i; term_at<R>(cf + " << off << ", d0r, d0i, d1r, d1i); const bool tk = "
                      << cond_expr(t.glb_ctrl, t.lane_ctrl) << ";";
                    if (uniform) {
                        o << " const bool sel = " << sel_expr(t.lane_sel, t.glb_sel)
                          << "; mul_if<R>(tk, f1r, f1i, sel ? d1r : d0r, sel ? d1i : d0i); }\n";
                    }
                    else {
                        o << " mul_if<R>(tk, f0r, f0i, d0r, d0i); mul_if<R>(tk, f1r, f1i, d1r, d1i); }\n";
                    }
                }
                if (!uniform) {
                    o << "    const Coef<R> fm {{f0r, f0i, 0, 0, 0, 0, f1r, f1i}};\n";
                }
            }
            else if (g.code >= TV_UNIFORM_BASE) {
                const size_t off = take();
                o << "    const Coef<R> m = coef_at<R>(cf + " << off << ");\n    const bool sel = " << sel_expr(g.lane_sel, g.glb_sel)
                  << ";\n    const R fr = sel ? m.m[6] : m.m[0], fi = sel ? m.m[7] : m.m[1];\n";
            }
            else if (ac.forms[gi].empty()) {
                const size_t off = take();
                o << "    const Coef<R> m = coef_at<R>(cf + " << off << ");\n";
            }
            if (g.pad != 0) {
                const TileGate& dg = tp.derivs[g.pad - 1];
                const size_t doff = take();
                o << "    {\n      const Coef<R> dm = coef_at<R>(cf + " << doff << ");\n      double acc = 0.0;\n"
                  << "      if (" << cond_expr(dg.glb_ctrl, dg.lane_ctrl) << ") {\n";
                if (dg.code >= TV_DGEN_BASE && dg.code < TV_DGEN_END) {
                    // an angle gate's fixed generator: its own code, no coefficients
                    const uint32_t pc = (dg.code - TV_DGEN_BASE) % 20;
                    o << "        acc = ddot_gen<R, " << (dg.code - TV_DGEN_BASE) / 20 << ", " << pc / 5 << ", " << pc % 5
                      << ">(v, w);\n";
                }
                else if (dg.code >= TV_UNIFORM_BASE) {
                    o << "        const bool dsel = " << sel_expr(dg.lane_sel, dg.glb_sel) << ";\n"
                      << "        acc = ddot_uniform<R, " << (dg.code - TV_UNIFORM_BASE)
                      << ">(v, w, dsel ? dm.m[6] : dm.m[0], dsel ? dm.m[7] : dm.m[1]);\n";
                }
                else {
                    o << "        acc = ddot<R, " << (dg.code / 5) % 4 << ", " << dg.code % 5 << ">(v, w, dm);\n";
                }
                if (staged) {
                    o << "      }\n      dp[" << pend.size() * 64 << "u + lane] = acc;\n    }\n";
                    pend.push_back(dg.pad);
                    if (pend.size() == STAGE) {
                        flush();
                    }
                }
                else {
                    o << "      }\n      acc = wsum(acc);\n      if ((tid & 63u) == 0u) {\n        partials[uint64_t {" << dg.pad
                      << "u} * n_part + col] = acc;\n      }\n    }\n";
                }
            }
            if (!ac.forms[gi].empty()) {
                // uncontrolled one-qubit gate, pivot divided out, on both vectors
                o << "  " << ac.forms[gi] << "  }\n";
                continue;
            }
            o << "    if (ok) { " << apply_to(g, "v") << " }\n";
            o << "    if (ok) { " << apply_to(g, "w") << " }\n  }\n";
        }
    }
    flush();
    if (std::any_of(ac.forms.begin(), ac.forms.end(), [](const std::string& f) { return !f.empty(); })) {
        // the product of the pivots taken out, last in the buffer, back into both vectors
        const size_t nf = ac.coefs.size();
        o << "  scale<R, 0>(v, static_cast<R>(cf[" << nf - 2 << "]), static_cast<R>(cf[" << nf - 1 << "]));\n"
          << "  scale<R, 0>(w, static_cast<R>(cf[" << nf - 2 << "]), static_cast<R>(cf[" << nf - 1 << "]));\n";
    }
    if (tp.out_layout != cur) {
        remap(tp.layouts[cur], tp.layouts[tp.out_layout]);
        cur = tp.out_layout;
    }
    io(tp.layouts[cur], false);
    o << "}\n";
    return o.str();
}

void launch_jit_adjoint_pass(kq_state* s, hipFunction_t fn, const TileProgram& tp, void* psi, void* lam,
                             const double* d_coefs, double* d_partials)
{
    if (tp.freepos.size() > 31 || tp.freepos.size() + tp.k != s->nvec) {
        throw InvalidArg {"jit adjoint pass: bad program shape"};
    }
    const double* cf = d_coefs;
    double* part = d_partials;
    const uint32_t n_blocks = tile_pass_blocks(tp);
    uint32_t n_part = n_blocks * tile_pass_waves(tp);
    void* args[] = {&psi, &lam, &cf, &part, &n_part};
    KQ_HIP_CHECK(hipModuleLaunchKernel(fn, n_blocks, 1, 1, 1u << tp.n_thr_bits, 1, 1, 0, s->stream, args, nullptr));
}

bool jit_available()
{
    JitCache& c = cache();
    if (c.available >= 0) {
        return c.available == 1;
    }
    int major = 0, minor = 0;
    c.available = hiprtcVersion(&major, &minor) == HIPRTC_SUCCESS ? 1 : 0;
    return c.available == 1;
}

namespace
{

struct Compiled
{
    std::vector<char> code;
    std::string log;
    bool ok = false;
    bool uncapped = false;   // compiled again without the register cap (the capped kernel spilled)
};

Compiled compile_one(const std::string& src)
{
    Compiled out;
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "kq_pass.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        out.log = "hiprtcCreateProgram failed";
        return out;
    }
    // the resource-usage remarks in the log tell whether a register cap made the kernel spill
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"};
    const hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(sizeof opts / sizeof opts[0]), opts);
    size_t n = 0;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
        out.log.resize(n);
        (void) hiprtcGetProgramLog(prog, out.log.data());
    }
    if (r == HIPRTC_SUCCESS && hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS) {
        out.code.resize(n);
        out.ok = hiprtcGetCode(prog, out.code.data()) == HIPRTC_SUCCESS;
    }
    (void) hiprtcDestroyProgram(&prog);
    return out;
}

}  // namespace

namespace
{

// ScratchSize [bytes/lane] of the kernel-resource-usage remark in a compile log (0 if absent)
size_t scratch_bytes(const std::string& log)
{
    const std::string key = "ScratchSize [bytes/lane]: ";
    const size_t at = log.find(key);
    return at == std::string::npos ? 0 : static_cast<size_t>(std::strtoull(log.c_str() + at + key.size(), nullptr, 10));
}

// compile sources[idx[j]] for every j on up to 16 host threads
std::vector<Compiled> compile_all(const std::vector<std::string>& sources, const std::vector<size_t>& idx)
{
    std::vector<Compiled> res(idx.size());
    std::atomic<size_t> next {0};
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>({idx.size(), hw, 16});
    std::vector<std::thread> pool;
    for (size_t w = 0; w < workers; ++w) {
        pool.emplace_back([&] {
            for (size_t j = next++; j < idx.size(); j = next++) {
                res[j] = compile_one(sources[idx[j]]);
                // a kernel that spills under the register cap is compiled again without the cap: private
                // (scratch) memory traffic costs more than the occupancy the cap buys
                const std::string& src = sources[idx[j]];
                const size_t cap = src.find(" __attribute__((amdgpu_waves_per_eu(");
                if (res[j].ok && cap != std::string::npos && scratch_bytes(res[j].log) > 0) {
                    const size_t end = src.find(")))", cap) + 3;   // __attribute__((amdgpu_waves_per_eu(N)))
                    Compiled uncapped = compile_one(src.substr(0, cap) + src.substr(end));
                    if (uncapped.ok) {
                        res[j] = std::move(uncapped);
                        res[j].uncapped = true;
                    }
                }
            }
        });
    }
    for (auto& t : pool) {
        t.join();
    }
    return res;
}

}  // namespace

// Kernels for the given sources (compiling the missing ones in parallel); null where compilation failed.
hipFunction_t jit_lookup(const std::string& source)
{
    JitCache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.kernels.find(source);
    if (it == c.kernels.end()) {
        return nullptr;
    }
    ++c.stats.hits;
    return it->second;
}

std::vector<hipFunction_t> jit_kernels(const std::vector<std::string>& sources)
{
    JitCache& c = cache();
    std::vector<hipFunction_t> out(sources.size(), nullptr);
    std::vector<size_t> missing;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        std::unordered_map<std::string, size_t> first;
        for (size_t i = 0; i < sources.size(); ++i) {
            auto it = c.kernels.find(sources[i]);
            if (it != c.kernels.end()) {
                out[i] = it->second;
                ++c.stats.hits;
            }
            else if (first.emplace(sources[i], i).second) {
                missing.push_back(i);
            }
        }
    }
    if (!missing.empty()) {
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<Compiled> res = compile_all(sources, missing);
        std::lock_guard<std::mutex> lk(c.mu);
        for (size_t j = 0; j < missing.size(); ++j) {
            if (!res[j].ok) {
                ++c.stats.failures;
                c.stats.last_log = res[j].log;
                continue;
            }
            hipModule_t mod = nullptr;
            hipFunction_t fn = nullptr;
            if (hipModuleLoadData(&mod, res[j].code.data()) != hipSuccess
                || hipModuleGetFunction(&fn, mod, sources[missing[j]].find(" kq_adjoint_pass(") != std::string::npos
                                                      ? "kq_adjoint_pass" : "kq_pass") != hipSuccess) {
                (void) hipGetLastError();
                ++c.stats.failures;
                c.stats.last_log = "module load failed";
                continue;
            }
            c.modules.push_back(mod);
            c.kernels.emplace(sources[missing[j]], fn);
            ++c.stats.compiled;
            c.stats.uncapped += res[j].uncapped ? 1 : 0;
        }
        c.stats.compile_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (size_t i = 0; i < sources.size(); ++i) {
            if (!out[i]) {
                auto it = c.kernels.find(sources[i]);
                out[i] = it == c.kernels.end() ? nullptr : it->second;
            }
        }
    }
    return out;
}

size_t jit_compile_only(const std::vector<std::string>& sources, std::string* log)
{
    std::vector<size_t> idx(sources.size());
    for (size_t i = 0; i < idx.size(); ++i) {
        idx[i] = i;
    }
    size_t failures = 0, uncapped = 0;
    for (const Compiled& c : compile_all(sources, idx)) {
        uncapped += c.uncapped ? 1 : 0;
        if (!c.ok) {
            ++failures;
            if (log) {
                *log = c.log;
            }
        }
    }
    std::lock_guard<std::mutex> lk(cache().mu);
    cache().stats.uncapped += uncapped;
    return failures;
}

void launch_jit_pass(kq_state* s, hipFunction_t fn, const TileProgram& tp, const double* d_coefs, const uint64_t* extra,
                     size_t n_extra)
{
    if (tp.freepos.size() > 31 || tp.freepos.size() + tp.k != s->nvec) {
        throw InvalidArg {"jit pass: bad program shape"};
    }
    const bool oop = !tp.out_freepos.empty();
    if (oop && (s->spare == nullptr || tp.out_freepos.size() != tp.freepos.size())) {
        throw InvalidArg {"jit pass: out-of-place pass without a destination buffer"};
    }
    if (n_extra > 2) {
        throw InvalidArg {"jit pass: too many arguments"};
    }
    void* amps = s->amps;
    void* dst = s->spare;
    const double* cf = d_coefs;
    uint32_t n_blocks = tile_pass_blocks(tp);
    uint64_t x[2] = {n_extra > 0 ? extra[0] : 0, n_extra > 1 ? extra[1] : 0};
    // the kernel's parameters: a, cf, n_blocks, [dst], then its input mode's values (basis, or zmask and zval)
    void* args[6] = {&amps, &cf, &n_blocks};
    size_t na = 3;
    if (oop) {
        args[na++] = &dst;
    }
    for (size_t i = 0; i < n_extra; ++i) {
        args[na++] = &x[i];
    }
    const JitTuning tune = jit_tuning();
    if (n_extra > 0 && tune.blocks_per_cu > 0) {
        throw InvalidArg {"jit pass: a pass from a basis state or over zero tiles has no persistent form"};
    }
    unsigned grid = n_blocks;
    if (tune.blocks_per_cu > 0) {
        grid = static_cast<unsigned>(std::min<uint64_t>(n_blocks, uint64_t {device_cu_count()} * tune.blocks_per_cu));
    }
    const unsigned threads = 1u << tp.n_thr_bits;
    KQ_HIP_CHECK(hipModuleLaunchKernel(fn, grid, 1, 1, threads, 1, 1, 0, s->stream, args, nullptr));
}

}  // namespace kq
