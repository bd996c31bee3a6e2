/*
 * synth.c -- synthetic growth-machine inputs (host set-up, not the hot path).
 *
 * SCORE builds its deformed structure from cuboids or EBSD columns
 * (SCORE_StructureSynthesis.cpp:355-528); no Voronoi path is reachable in the
 * reference (POISSONVORONOI_DEFMS is parsed at SCORE_Kernel.cpp:453 but never
 * dispatched at :6155-6165).  The workloads here are Voronoi tessellations, so
 * this file writes the cell grid the reference would hold after
 * solve_REPLACE_CA_STRUCTURE, plus the pools, from a seeded generator.
 *
 * Everything is deterministic for a given seed and independent of the number
 * of OpenMP threads: each voxel's owner is computed on its own, ties between
 * equidistant seeds go to the lower seed index.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* splitmix64 */
static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform double in [0,1) with 53 random bits */
static double u01(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* uniform integer in [0,n) without modulo bias */
static uint64_t below(uint64_t *s, uint64_t n) {
    uint64_t lim = UINT64_MAX - (UINT64_MAX % n);
    uint64_t r;
    do { r = sm64(s); } while (r >= lim);
    return r % n;
}

/* Uniform random unit quaternions (Shoemake, Graphics Gems III, 1992). */
void synth_quaternions(uint64_t seed, uint32_t n, double *q) {
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; i++) {
        double u1 = u01(&s), u2 = u01(&s), u3 = u01(&s);
        double a = sqrt(1.0 - u1), b = sqrt(u1);
        double t1 = 2.0 * 3.14159265358979323846 * u2, t2 = 2.0 * 3.14159265358979323846 * u3;
        q[4 * (size_t)i + 0] = b * cos(t2);
        q[4 * (size_t)i + 1] = a * sin(t1);
        q[4 * (size_t)i + 2] = a * cos(t1);
        q[4 * (size_t)i + 3] = b * sin(t2);
    }
}

/* n values 10^(lo + (hi-lo) U) */
void synth_loguniform(uint64_t seed, uint32_t n, double lo, double hi, double *out) {
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; i++) out[i] = pow(10.0, lo + (hi - lo) * u01(&s));
}

/* n values exp(N(mu, sigma)) by Box-Muller */
void synth_lognormal(uint64_t seed, uint32_t n, double mu, double sigma, double *out) {
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; i++) {
        double u1 = u01(&s), u2 = u01(&s);
        if (u1 < 1e-300) u1 = 1e-300;
        double g = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
        out[i] = exp(mu + sigma * g);
    }
}

/* n uniform points in [0,nx)x[0,ny)x[0,nz) */
void synth_points(uint64_t seed, uint32_t n, uint32_t nx, uint32_t ny, uint32_t nz, double *xyz) {
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; i++) {
        xyz[3 * (size_t)i + 0] = u01(&s) * nx;
        xyz[3 * (size_t)i + 1] = u01(&s) * ny;
        xyz[3 * (size_t)i + 2] = u01(&s) * nz;
    }
}

/* n distinct uniform sites in [0, N); returns 0 or -1 when n > N or on OOM */
int synth_distinct_sites(uint64_t seed, uint32_t n, uint64_t N, uint32_t *sites) {
    if ((uint64_t)n > N) return -1;
    uint64_t words = (N + 63) / 64;
    uint64_t *bits = (uint64_t *)calloc(words, sizeof(uint64_t));
    if (!bits) return -1;
    uint64_t s = seed;
    for (uint32_t i = 0; i < n;) {
        uint64_t w = below(&s, N);
        uint64_t m = 1ull << (w & 63);
        if (bits[w >> 6] & m) continue;
        bits[w >> 6] |= m;
        sites[i++] = (uint32_t)w;
    }
    free(bits);
    return 0;
}

/* Whether the cell at linear index w has a 26-neighbour (periodic in x and y,
 * open in z) in another grain. */
static int on_boundary(const uint32_t *grid, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t w) {
    const uint64_t nxy = (uint64_t)nx * ny;
    const uint32_t z = (uint32_t)(w / nxy), y = (uint32_t)((w % nxy) / nx), x = (uint32_t)(w % nx);
    const uint32_t g = grid[w];
    for (int dz = -1; dz <= 1; dz++) {
        const int zz = (int)z + dz;
        if (zz < 0 || zz >= (int)nz) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const uint32_t yy = (uint32_t)(((int)y + dy + (int)ny) % (int)ny);
            for (int dx = -1; dx <= 1; dx++) {
                const uint32_t xx = (uint32_t)(((int)x + dx + (int)nx) % (int)nx);
                if (grid[xx + (uint64_t)yy * nx + (uint64_t)zz * nxy] != g) return 1;
            }
        }
    }
    return 0;
}

/* n distinct sites drawn one after another without replacement, a cell next to
 * a grain boundary (a 26-neighbour in another grain) `bias` times as likely as
 * an interior one: a uniform candidate is kept with probability 1 on a
 * boundary, 1/bias inside (bias >= 1).  Returns 0, or -1 when n exceeds the
 * cells or on OOM. */
int synth_distinct_sites_gb(uint64_t seed, uint32_t n, const uint32_t *grid, uint32_t nx, uint32_t ny, uint32_t nz,
                            double bias, uint32_t *sites) {
    const uint64_t N = (uint64_t)nx * ny * nz;
    if ((uint64_t)n > N || !(bias >= 1.0)) return -1;
    uint64_t *bits = (uint64_t *)calloc((N + 63) / 64, sizeof(uint64_t));
    if (!bits) return -1;
    uint64_t s = seed;
    for (uint32_t i = 0; i < n;) {
        const uint64_t w = below(&s, N);
        const double u = u01(&s);
        const uint64_t m = 1ull << (w & 63);
        if (bits[w >> 6] & m) continue;
        if (!on_boundary(grid, nx, ny, nz, w) && u * bias >= 1.0) continue;
        bits[w >> 6] |= m;
        sites[i++] = (uint32_t)w;
    }
    free(bits);
    return 0;
}

/*
 * Nearest-seed labelling of the nx*ny*nz voxel centres (i+0.5, j+0.5, k+0.5).
 * Metric: ((dx/sx)^2 + (dy/sy)^2 + (dz/sz)^2), so sx > 1 stretches grains along
 * x.  Periodic images in x and y when periodic_xy != 0 (SCORE wraps x and y,
 * not z: SCORE_Kernel.cpp:4362-4377).  Seeds are bucketed on a grid and the
 * search grows shell by shell until no unvisited bucket can hold a closer seed.
 */
int synth_voronoi(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nseeds, const double *xyz, double sx, double sy,
                  double sz, int periodic_xy, uint32_t *grid) {
    if (nseeds == 0) return -1;
    double vol = (double)nx * ny * nz;
    double edge = cbrt(vol / nseeds * 2.0); /* ~2 seeds per bucket */
    int bx = (int)ceil(nx / (edge * sx)); if (bx < 1) bx = 1; if (bx > 512) bx = 512;
    int by = (int)ceil(ny / (edge * sy)); if (by < 1) by = 1; if (by > 512) by = 512;
    int bz = (int)ceil(nz / (edge * sz)); if (bz < 1) bz = 1; if (bz > 512) bz = 512;
    double wx = (double)nx / bx, wy = (double)ny / by, wz = (double)nz / bz;
    size_t nb = (size_t)bx * by * bz;
    uint32_t *start = (uint32_t *)calloc(nb + 1, sizeof(uint32_t));
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * nseeds);
    uint32_t *bid = (uint32_t *)malloc(sizeof(uint32_t) * nseeds);
    if (!start || !order || !bid) { free(start); free(order); free(bid); return -1; }
    for (uint32_t i = 0; i < nseeds; i++) {
        int ix = (int)(xyz[3 * (size_t)i] / wx), iy = (int)(xyz[3 * (size_t)i + 1] / wy), iz = (int)(xyz[3 * (size_t)i + 2] / wz);
        if (ix >= bx) ix = bx - 1; if (iy >= by) iy = by - 1; if (iz >= bz) iz = bz - 1;
        bid[i] = (uint32_t)(ix + bx * (iy + by * iz));
        start[bid[i] + 1]++;
    }
    for (size_t b = 0; b < nb; b++) start[b + 1] += start[b];
    {
        uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * nb);
        if (!fill) { free(start); free(order); free(bid); return -1; }
        memcpy(fill, start, sizeof(uint32_t) * nb);
        for (uint32_t i = 0; i < nseeds; i++) order[fill[bid[i]]++] = i; /* ascending seed index per bucket */
        free(fill);
    }
    /* smallest scaled distance from any point of a bucket to a bucket r shells away */
    double shell = fmin(fmin(wx / sx, wy / sy), wz / sz);
    int rmax = bx; if (by > rmax) rmax = by; if (bz > rmax) rmax = bz;

#pragma omp parallel for schedule(dynamic, 4)
    for (long long k = 0; k < (long long)nz; k++) {
        for (uint32_t j = 0; j < ny; j++) {
            for (uint32_t i = 0; i < nx; i++) {
                double px = i + 0.5, py = j + 0.5, pz = k + 0.5;
                int cx = (int)(px / wx), cy = (int)(py / wy), cz = (int)(pz / wz);
                if (cx >= bx) cx = bx - 1; if (cy >= by) cy = by - 1; if (cz >= bz) cz = bz - 1;
                double best = INFINITY;
                uint32_t bestid = UINT32_MAX;
                for (int r = 0; r <= rmax; r++) {
                    /* every bucket at Chebyshev distance r in bucket space */
                    for (int dz = -r; dz <= r; dz++) {
                        int zz = cz + dz;
                        if (zz < 0 || zz >= bz) continue;
                        for (int dy = -r; dy <= r; dy++) {
                            int yy = cy + dy;
                            if (!periodic_xy && (yy < 0 || yy >= by)) continue;
                            if (periodic_xy && (dy < -by / 2 - 1 || dy > by / 2 + 1)) continue;
                            for (int dx = -r; dx <= r; dx++) {
                                if (abs(dx) != r && abs(dy) != r && abs(dz) != r) continue;
                                int xx = cx + dx;
                                if (!periodic_xy && (xx < 0 || xx >= bx)) continue;
                                if (periodic_xy && (dx < -bx / 2 - 1 || dx > bx / 2 + 1)) continue;
                                int wxx = ((xx % bx) + bx) % bx, wyy = ((yy % by) + by) % by;
                                size_t b = (size_t)wxx + (size_t)bx * ((size_t)wyy + (size_t)by * zz);
                                for (uint32_t t = start[b]; t < start[b + 1]; t++) {
                                    uint32_t s = order[t];
                                    double ddx = px - xyz[3 * (size_t)s], ddy = py - xyz[3 * (size_t)s + 1];
                                    double ddz = pz - xyz[3 * (size_t)s + 2];
                                    if (periodic_xy) {
                                        if (ddx > 0.5 * nx) ddx -= nx; else if (ddx < -0.5 * nx) ddx += nx;
                                        if (ddy > 0.5 * ny) ddy -= ny; else if (ddy < -0.5 * ny) ddy += ny;
                                    }
                                    ddx /= sx; ddy /= sy; ddz /= sz;
                                    double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                    if (d2 < best || (d2 == best && s < bestid)) { best = d2; bestid = s; }
                                }
                            }
                        }
                    }
                    /* any bucket beyond shell r is at least r*shell away */
                    double lim = r * shell;
                    if (bestid != UINT32_MAX && lim * lim > best) break;
                }
                grid[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)] = bestid;
            }
        }
    }
    free(start); free(order); free(bid);
    return 0;
}

/* Brute-force labelling with the same metric and tie rule, for checking synth_voronoi. */
int synth_voronoi_bruteforce(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nseeds, const double *xyz, double sx,
                             double sy, double sz, int periodic_xy, uint32_t *grid) {
    for (uint32_t k = 0; k < nz; k++)
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++) {
                double px = i + 0.5, py = j + 0.5, pz = k + 0.5, best = INFINITY;
                uint32_t bestid = UINT32_MAX;
                for (uint32_t s = 0; s < nseeds; s++) {
                    double ddx = px - xyz[3 * (size_t)s], ddy = py - xyz[3 * (size_t)s + 1], ddz = pz - xyz[3 * (size_t)s + 2];
                    if (periodic_xy) {
                        if (ddx > 0.5 * nx) ddx -= nx; else if (ddx < -0.5 * nx) ddx += nx;
                        if (ddy > 0.5 * ny) ddy -= ny; else if (ddy < -0.5 * ny) ddy += ny;
                    }
                    ddx /= sx; ddy /= sy; ddz /= sz;
                    double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (d2 < best) { best = d2; bestid = s; }
                }
                grid[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)] = bestid;
            }
    return 0;
}
