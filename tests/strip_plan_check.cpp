// Host-side check of the stage kernels' work decomposition (strip_plan, segment_of in
// wenofv_kernels.h, the functions the kernels call): over many grid shapes, rows per strip
// and both stencil halos, every cell is owned by exactly one lane of one wavefront; each
// segment lies inside its wavefront; and every lane an owned cell depends on (the x-stencil
// lanes of the reconstructions on both sides of its faces, the lanes solving its east and
// west faces) sits in the same segment, on the adjacent column, with the same rows.
// Also the autotuner's occupancy candidates and the XCD-aware strip and batch item orders.
// Prints "ok <cases>" or the first failure and exits 1.
#include <cstdio>
#include <vector>

#include "wenofv_host.h"
#include "wenofv_kernels.h"

using wfv::Segment;
using wfv::StripPlan;

static int fail(const char *what, int nx, int ny, int ty, int halo, int strip, int lane) {
    std::printf("FAIL %s nx=%d ny=%d ty=%d halo=%d strip=%d lane=%d\n", what, nx, ny, ty, halo, strip, lane);
    return 1;
}

static int check(int nx, int ny, int ty, int halo) {
    const int L = wfv::WAVE_LANES, R = halo - 1, SW = L - 2 * halo;
    const StripPlan sp = wfv::strip_plan(nx, ny, ty, halo);
    std::vector<int> count((size_t)nx * ny, 0);
    for (int strip = 0; strip < sp.total; ++strip) {
        Segment seg[64];
        for (int lane = 0; lane < L; ++lane) seg[lane] = wfv::segment_of(sp, strip, lane, ty, ny, halo);
        for (int lane = 0; lane < L; ++lane) {
            const Segment &g = seg[lane];
            if (g.segw < 1 || g.segw > L || g.j < 0 || g.j >= g.segw || lane - g.j + g.segw > L)
                return fail("segment outside the wavefront", nx, ny, ty, halo, strip, lane);
            // a segment with rows is wide enough for its stencils; the kernel's stencil
            // (recon) and face lanes of any segment read only lanes of that segment
            if (g.y1 > g.y0 && g.segw < 2 * halo + 1) return fail("narrow segment with rows", nx, ny, ty, halo, strip, lane);
            const bool recon = g.j >= R && g.j <= g.segw - 1 - R, xface = g.j >= R && g.j <= g.segw - 2 - R;
            if ((recon && (g.j - R < 0 || g.j + R > g.segw - 1)) || (xface && g.j + 1 > g.segw - 1))
                return fail("stencil lane reads outside its segment", nx, ny, ty, halo, strip, lane);
            if (g.y0 < 0 || g.y0 > g.y1 || g.y1 > ny || g.y1 - g.y0 > g.nrow_w)
                return fail("rows", nx, ny, ty, halo, strip, lane);
            const int xc = g.x0 - halo + g.j;
            const bool owned = g.j >= halo && g.j < halo + SW && xc < nx && g.y1 > g.y0;
            if (!owned) continue;
            // lanes j-1 and j solve the west and east faces (xface lanes); lanes j-1..j+1
            // reconstruct (recon lanes) from x-stencils j'-R..j'+R
            for (int d = -1; d <= 1; ++d) {
                const int jj = g.j + d;
                if (jj - R < 0 || jj + R > g.segw - 1) return fail("stencil leaves the segment", nx, ny, ty, halo, strip, lane);
                if (d <= 0 && !(jj >= R && jj <= g.segw - 2 - R)) return fail("face lane", nx, ny, ty, halo, strip, lane);
            }
            for (int d = -R - 1; d <= R + 1; ++d) {
                const Segment &n = seg[lane + d];
                if (n.y0 != g.y0 || n.y1 != g.y1 || n.x0 - halo + n.j != xc + d)
                    return fail("neighbour lane", nx, ny, ty, halo, strip, lane);
            }
            for (int y = g.y0; y < g.y1; ++y) ++count[(size_t)y * nx + xc];
        }
    }
    for (size_t i = 0; i < count.size(); ++i)
        if (count[i] != 1) return fail("cell not owned exactly once", nx, ny, ty, halo, (int)(i / nx), (int)(i % nx));
    return 0;
}

// The autotuner's occupancy candidates (occupancy_candidates) on a 1024 x 512 grid of the
// 5-point schemes (halo 3) and a 256-CU device: 18 strips across; 10 rows fill one round of
// one wavefront per SIMD (936 <= 1024 wavefronts; 9 rows would need 1,026), 5 rows two such
// rounds or one of two per SIMD (1,854 <= 2,048), 4 and 3 rows three and four rounds
// (2,304 <= 3,072, 3,078 <= 4,096), 2 rows three rounds of two per SIMD (4,608 <= 6,144).
static int check_occupancy() {
    int out[8];
    const int n = wfv::occupancy_candidates(1024, 512, 3, 256, out);
    const int want[8] = {10, 5, 4, 3, 5, 3, 2, 2};
    if (n != 8) return fail("occupancy candidates: count", 1024, 512, n, 3, 0, 0);
    for (int k = 0; k < 8; ++k)
        if (out[k] != want[k]) return fail("occupancy candidates: value", 1024, 512, out[k], 3, k, want[k]);
    // every candidate's wavefronts fit its rounds, and one row fewer would not
    const int caps[8] = {1024, 2048, 3072, 4096, 2048, 4096, 6144, 8192};
    for (int k = 0; k < 8; ++k) {
        if (wfv::strip_plan(1024, 512, out[k], 3).total > caps[k]) return fail("occupancy: over", 1024, 512, out[k], 3, k, 0);
        if (out[k] > 1 && wfv::strip_plan(1024, 512, out[k] - 1, 3).total <= caps[k])
            return fail("occupancy: not the fewest rows", 1024, 512, out[k], 3, k, 0);
    }
    return 0;
}

// The XCD-aware strip order (xcd_strip): a bijection of [0, n) under which the workgroups of
// each XCD (b mod 8) take one contiguous run of strips in dispatch order, the runs' lengths
// differing by at most one.
static int check_xcd_order() {
    for (int n = 1; n <= 20000; n += n < 600 ? 1 : 37) {
        std::vector<int> seen(n, 0), first(8, -1), last(8, -1), len(8, 0);
        for (int b = 0; b < n; ++b) {
            const int s = wfv::xcd_strip(b, n), x = b % 8;
            if (s < 0 || s >= n || seen[s]++) return fail("xcd order: not a bijection", n, 0, 0, 0, b, s);
            if (last[x] >= 0 && s != last[x] + 1) return fail("xcd order: run not contiguous", n, 0, 0, 0, b, s);
            if (first[x] < 0) first[x] = s;
            last[x] = s;
            ++len[x];
        }
        for (int x = 0; x < 8; ++x)
            if (len[x] != n / 8 + (x < n % 8 ? 1 : 0)) return fail("xcd order: run length", n, 0, 0, 0, x, len[x]);
    }
    return 0;
}

// The batch's XCD-aware item order (batch_item_order) over seeded random instance lists: every
// (instance, strip) item exactly once; each XCD's items (positions x, x+8, ...) a prefix
// without gaps, in instance-priority then strip order; the XCDs' work within one run of
// the largest per-strip work of the smallest load.
static int check_batch_order() {
    unsigned seed = 12345u;
    auto rnd = [&](int n) { seed = seed * 1664525u + 1013904223u; return (int)((seed >> 8) % (unsigned)n); };
    const int rows_set[] = {1, 2, 4, 12, 48, 64, 96, 128};
    for (int trial = 0; trial < 200; ++trial) {
        const int n = 1 + rnd(300), chunk = 1 + rnd(32);
        std::vector<int> nstrip(n), rows(n), prio(n);
        for (int i = 0; i < n; ++i) {
            nstrip[i] = 1 + rnd(trial % 3 == 0 ? 20 : 600);
            rows[i] = rows_set[rnd(8)];
            prio[i] = i;
        }
        std::stable_sort(prio.begin(), prio.end(), [&](int a, int b) { return rows[a] > rows[b]; });
        std::vector<int> rank(n);
        for (int r = 0; r < n; ++r) rank[prio[r]] = r;
        std::vector<int> ii, ss;
        const int P = wfv::batch_item_order(prio, nstrip, rows, chunk, ii, ss);
        if (P % 8 != 0 || (int)ii.size() != P || (int)ss.size() != P) return fail("batch order: size", n, chunk, P, 0, trial, 0);
        std::vector<std::vector<char>> seen(n);
        for (int i = 0; i < n; ++i) seen[i].assign(nstrip[i], 0);
        long long total = 0;
        double load[8] = {}, maxw = 0;
        for (int x = 0; x < 8; ++x) {
            bool ended = false;
            int prank = -1, ps = -1;
            for (int p = x; p < P; p += 8) {
                const int i = ii[p], st = ss[p];
                if (i < 0) { ended = true; continue; }
                if (ended) return fail("batch order: gap before an item", n, chunk, P, 0, trial, p);
                if (i >= n || st < 0 || st >= nstrip[i] || seen[i][st]++) return fail("batch order: item", n, chunk, P, 0, trial, p);
                if (rank[i] < prank || (rank[i] == prank && st <= ps)) return fail("batch order: order", n, chunk, P, 0, trial, p);
                prank = rank[i];
                ps = st;
                load[x] += rows[i] + 1.0;
                maxw = rows[i] + 1.0 > maxw ? rows[i] + 1.0 : maxw;
                ++total;
            }
        }
        long long want = 0;
        for (int i = 0; i < n; ++i) want += nstrip[i];
        if (total != want) return fail("batch order: count", n, chunk, P, 0, trial, (int)total);
        double lo = load[0], hi = load[0];
        for (int x = 1; x < 8; ++x) {
            lo = load[x] < lo ? load[x] : lo;
            hi = load[x] > hi ? load[x] : hi;
        }
        if (hi - lo > chunk * maxw + 1e-9) return fail("batch order: balance", n, chunk, P, 0, trial, (int)(hi - lo));
    }
    return 0;
}

int main() {
    if (check_occupancy()) return 1;
    if (check_xcd_order()) return 1;
    if (check_batch_order()) return 1;
    const int nxs[] = {1, 5, 6, 7, 16, 20, 25, 26, 27, 31, 32, 50, 56, 57, 58, 59, 63, 64, 70, 100,
                       112, 115, 116, 117, 128, 170, 174, 200, 256, 300, 512, 1000, 1024};
    const int nys[] = {1, 2, 3, 7, 16, 33, 64, 100, 128, 257};
    // every rows-per-strip value up to 40 (the autotuner's occupancy candidates can be any
    // value) and the larger ones it tries or may pick
    std::vector<int> tys;
    for (int t = 1; t <= 40; ++t) tys.push_back(t);
    for (int t : {44, 48, 50, 56, 60, 64, 74, 76, 80, 96, 100, 112, 127, 128}) tys.push_back(t);
    int cases = 0;
    for (int halo : {3, 4})
        for (int nx : nxs)
            for (int ny : nys)
                for (int ty : tys) {
                    if (check(nx, ny, ty, halo)) return 1;
                    ++cases;
                }
    std::printf("ok %d\n", cases);
    return 0;
}
