// The flux-word layout of the spherical engine (ffidx, ffown), checked on
// the CPU for both block shapes (bz = 16, 1): every face the flux kernels
// form has its own in-range slot, and every read -- the EMF stencils of the
// face update over all 64 combinations of shared sides (with the neighbour
// remapping of the EMF reads), the flux divergence of the zone update, the
// line kernel -- lands on a slot that is written.  Built and run by
// tests/test_layout.py.
// The face sets mirror the engine's face lists (build_item_lists in
// sph_device.hip), the reads emf_x/y/z with Fsh and cell_hydro.
#include <cstdio>
#include <vector>
#include <set>
#include <cstdlib>
#include "sph_common.hpp"
using namespace sph;
int main() {
    for (int bz : {16, 1}) {
        Lay L = make_lay(bz);
        const int n[3] = {BX, BY, bz};
        int bad = 0;
        // written slots per direction: own and per side
        std::vector<int> owner[3];
        for (int d = 0; d < 3; d++) owner[d].assign(L.nff, -1);  // -1 unwritten, 0 own, 1+sd side
        for (int d = 0; d < 3; d++) {
            int lo[3], hi[3];
            for (int e = 0; e < 3; e++) { lo[e] = e == d ? 0 : -1; hi[e] = n[e]; }
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        const int c[3] = {i, j, k};
                        int out = 0, sd = -1;
                        for (int e = 0; e < 3; e++)
                            if (e != d && (c[e] < 0 || c[e] >= n[e])) { out++; sd = 2 * e + (c[e] < 0 ? 0 : 1); }
                        if (out == 2) continue;
                        const int x = ffidx(d, i, j, k, bz);
                        if (x < 0 || x >= L.nff) { printf("bz %d OOB write d%d %d %d %d -> %d\n", bz, d, i, j, k, x); bad++; continue; }
                        if (owner[d][x] != -1) { printf("bz %d collision d%d %d %d %d\n", bz, d, i, j, k); bad++; }
                        owner[d][x] = out == 0 ? 0 : 1 + sd;
                        if (out == 0 && ffown(d, i, j, k, bz) != x) { printf("ffown mismatch\n"); bad++; }
                    }
        }
        // reads: Fsh with every combination of shared sides
        auto check_read = [&](int d, int i, int j, int k, int shmask) {
            // remap as Fsh
            int ii = i, jj = j, kk = k;
            bool remapped = false;
            if (d != 0) { if (ii < 0 && (shmask & 1)) { ii += BX; remapped = true; } else if (ii >= BX && (shmask & 2)) { ii -= BX; remapped = true; } }
            if (d != 1) { if (jj < 0 && (shmask & 4)) { jj += BY; remapped = true; } else if (jj >= BY && (shmask & 8)) { jj -= BY; remapped = true; } }
            if (d != 2) { if (kk < 0 && (shmask & 16)) { kk += bz; remapped = true; } else if (kk >= bz && (shmask & 32)) { kk -= bz; remapped = true; } }
            const int c[3] = {ii, jj, kk};
            int out = 0;
            for (int e = 0; e < 3; e++) if (e != d && (c[e] < 0 || c[e] >= n[e])) out++;
            if (c[d] < 0 || c[d] > n[d]) { printf("bz %d along-d out of range d%d %d %d %d\n", bz, d, ii, jj, kk); bad++; return; }
            if (out == 2) { printf("bz %d corner read d%d %d %d %d\n", bz, d, ii, jj, kk); bad++; return; }
            const int x = ffidx(d, ii, jj, kk, bz);
            if (x < 0 || x >= L.nff) { printf("bz %d OOB read\n", bz); bad++; return; }
            if (owner[d][x] < 0) { printf("bz %d read of unwritten slot\n", bz); bad++; return; }
            if (remapped && owner[d][x] != 0) { printf("bz %d remapped read not own\n", bz); bad++; }
            // a side slot read must be on a side the block does not share (it computes it)
            if (!remapped && owner[d][x] > 0) {
                const int sd = owner[d][x] - 1;
                if (shmask & (1 << sd)) { printf("bz %d read of a shared side's slot\n", bz); bad++; }
            }
        };
        long nreads = 0;
        for (int sh = 0; sh < 64; sh++) {
            // emf_z(i, j, k): i, j in [0, 16], k in [0, bz)
            for (int k = 0; k < bz; k++) for (int j = 0; j <= BY; j++) for (int i = 0; i <= BX; i++) {
                check_read(1, i, j, k, sh); check_read(1, i - 1, j, k, sh); check_read(0, i, j, k, sh); check_read(0, i, j - 1, k, sh); nreads += 4; }
            // emf_y(i, j, k): i in [0, 16], j in [0, 16), k in [0, bz]
            for (int k = 0; k <= bz; k++) for (int j = 0; j < BY; j++) for (int i = 0; i <= BX; i++) {
                check_read(0, i, j, k, sh); check_read(0, i, j, k - 1, sh); check_read(2, i, j, k, sh); check_read(2, i - 1, j, k, sh); nreads += 4; }
            // emf_x(i, j, k): i in [0, 16), j in [0, 16], k in [0, bz]
            for (int k = 0; k <= bz; k++) for (int j = 0; j <= BY; j++) for (int i = 0; i < BX; i++) {
                check_read(2, i, j, k, sh); check_read(2, i, j - 1, k, sh); check_read(1, i, j, k, sh); check_read(1, i, j, k - 1, sh); nreads += 4; }
        }
        // update reads (own faces through ffown)
        for (int k = 0; k < bz; k++) for (int j = 0; j < BY; j++) for (int i = 0; i < BX; i++) {
            int f[6][4] = {{0, i, j, k}, {0, i + 1, j, k}, {1, i, j, k}, {1, i, j + 1, k}, {2, i, j, k}, {2, i, j, k + 1}};
            for (auto& q : f) {
                const int x = ffown(q[0], q[1], q[2], q[3], bz);
                if (x < 0 || x >= L.nff || owner[q[0]][x] != 0 || x != ffidx(q[0], q[1], q[2], q[3], bz)) { printf("bz %d update read bad\n", bz); bad++; }
            }
        }
        // pencil lines: d = 1 lines (i, k), d = 2 lines (i, j), faces 0..n_d
        for (int k = -1; k <= bz; k++) for (int i = -1; i <= BX; i++) {
            const int oi = i < 0 || i >= BX, ok = k < 0 || k >= bz; if (oi && ok) continue;
            for (int f = 0; f <= BY; f++) { const int x = ffidx(1, i, f, k, bz); if (x < 0 || x >= L.nff || owner[1][x] < 0) { printf("pencil1 bad\n"); bad++; } } }
        for (int j = -1; j <= BY; j++) for (int i = -1; i <= BX; i++) {
            const int oi = i < 0 || i >= BX, oj = j < 0 || j >= BY; if (oi && oj) continue;
            for (int f = 0; f <= bz; f++) { const int x = ffidx(2, i, j, f, bz); if (x < 0 || x >= L.nff || owner[2][x] < 0) { printf("pencil2 bad\n"); bad++; } } }
        int unwritten = 0;
        for (int d = 0; d < 3; d++) for (int x = 0; x < L.nff; x++) if (owner[d][x] < 0) unwritten++;
        printf("bz %d: nff %d (nf %d), reads checked %ld, slots never written %d, errors %d\n", bz, L.nff, L.nf, nreads, unwritten, bad);
        if (bad) return 1;
    }
    return 0;
}
