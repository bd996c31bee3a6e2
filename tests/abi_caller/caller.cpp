// caller.cpp -- a C++ program that uses the growth machine the way SCORE's
// adapter would (caHdl::solve_RXGROWTH replaced by the C ABI of score_gm.h),
// compiled against include/score_gm.h and include/score_io.h with g++ -std=c++11
// (SCORE's language level) and linked to libscore_gm.so and libscore_io.so.
//
//   score_gm_caller <input.bin> <outdir>
//
// input.bin holds one domain and the few settings that vary (little-endian):
//   uint32 nx, ny, nz; double cellsize; uint32 n_ori, n_defg, n_rxg, n_log;
//   double quat[4 n_ori]; uint32 defg_ori[n_defg]; double rho[n_defg];
//   uint32 rxg_ori[n_rxg]; uint32 rxg_site[n_rxg]; uint32 cellgrid[N];
//   double log_at[n_log]; double T, m0, hact_J, xmax
// The program runs the domain to completion and writes into outdir what SCORE
// reads back after solve_RXGROWTH: result.bin, counts.bin, map.bin (1 =
// infected, else 2+gid), traj.bin, glog.bin, glogcounts.bin, status.bin, and
// the grain map as SCORE's raw stack, HDF5 dataset and XDMF file.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "score_gm.h"
#include "score_io.h"

namespace {

bool rd(FILE *f, void *p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

bool wr(const std::string &path, const void *p, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = n == 0 || fwrite(p, 1, n, f) == n;
    return fclose(f) == 0 && ok;
}

int die(const char *what, int rc) {
    fprintf(stderr, "caller: %s: %d (%s)\n", what, rc, score_gm_strerror(rc));
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin outdir\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t geo[3], cnt[4];
    double cellsize;
    if (!rd(f, geo, sizeof geo) || !rd(f, &cellsize, 8) || !rd(f, cnt, sizeof cnt)) return 2;
    const uint32_t nx = geo[0], ny = geo[1], nz = geo[2];
    const uint32_t n_ori = cnt[0], n_defg = cnt[1], n_rxg = cnt[2], n_log = cnt[3];
    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> quat(4 * (size_t)n_ori), rho(n_defg), log_at(n_log + 1);
    std::vector<uint32_t> dori(n_defg), rori(n_rxg + 1), site(n_rxg + 1), grid(N);
    double tail[4];
    if (!rd(f, quat.data(), 8 * quat.size()) || !rd(f, dori.data(), 4 * (size_t)n_defg) ||
        !rd(f, rho.data(), 8 * (size_t)n_defg) || !rd(f, rori.data(), 4 * (size_t)n_rxg) ||
        !rd(f, site.data(), 4 * (size_t)n_rxg) || !rd(f, grid.data(), 4 * N) || !rd(f, log_at.data(), 8 * (size_t)n_log) ||
        !rd(f, tail, sizeof tail))
        return 2;
    fclose(f);
    const std::string out = argv[2];

    // the adapter's translation of caHdl state (a Fe / Rollett-Holm deck)
    score_gm_domain dom;
    memset(&dom, 0, sizeof dom);
    dom.nx = nx; dom.ny = ny; dom.nz = nz; dom.cellsize = cellsize;
    dom.cellgrid = grid.data();
    dom.n_ori = n_ori; dom.ori_quat = quat.data();
    dom.n_defg = n_defg; dom.defg_ori = dori.data(); dom.defg_rho = rho.data();
    dom.n_rxg = n_rxg; dom.rxg_ori = rori.data(); dom.rxg_site = site.data();
    dom.rxg_step = NULL;  // site saturation
    const double sched_t[2] = {0.0, 3.6e5}, sched_T[2] = {tail[0], tail[0]};
    score_gm_params prm;
    memset(&prm, 0, sizeof prm);
    prm.mobility_model = SCORE_GM_MOB_ROLLETTHOLM;
    prm.rh_hagb_m0 = tail[1]; prm.rh_hagb_hact = tail[2];
    prm.rh_cut = 0.9; prm.rh_trans = 5.0; prm.rh_exponent = 9.0;
    prm.lagb_m0 = 0.3; prm.hagb_m0 = 3.0; prm.gs_m0 = 3.0;
    prm.lagb_hact = prm.hagb_hact = prm.gs_hact = 1.2 * 1.602176565e-19;
    prm.G0 = 6.92e10; prm.b0 = 2.48e-10;
    prm.n_sched = 2; prm.sched_time = sched_t; prm.sched_temp = sched_T;
    prm.maxfill = 0.1; prm.xmax = tail[3]; prm.tmax = 3.6e5; prm.nmax = 100000;
    prm.rel_cell_caching = 0.2; prm.rel_cell_recaching = 1.0;
    prm.n_grainlog = n_log; prm.grainlog_at = log_at.data();

    if (score_gm_abi_version() != SCORE_GM_ABI_VERSION) return die("ABI version", score_gm_abi_version());
    score_gm_handle *h = NULL;
    int rc = score_gm_create(&dom, &prm, 0, &h);
    if (rc) return die("score_gm_create", rc);
    score_gm_result res;
    if ((rc = score_gm_run(h, &res))) return die("score_gm_run", rc);

    // tsimend, stepsimend, t, X, Xcells, step, nmynuclei as the adapter writes them back
    double rs[4] = {res.t, res.X, res.xcells, res.pmax};
    uint32_t ri[4] = {res.step, (uint32_t)res.completed, res.nuclei_placed, 0};
    uint64_t sv = res.sum_sv;
    std::vector<unsigned char> rb(sizeof rs + sizeof ri + sizeof sv);
    memcpy(rb.data(), rs, sizeof rs);
    memcpy(rb.data() + sizeof rs, ri, sizeof ri);
    memcpy(rb.data() + sizeof rs + sizeof ri, &sv, sizeof sv);
    if (!wr(out + "/result.bin", rb.data(), rb.size())) return 3;

    std::vector<uint32_t> counts((size_t)n_defg + n_rxg + 1);
    if ((rc = score_gm_get_counts(h, counts.data(), counts.data() + n_defg))) return die("score_gm_get_counts", rc);
    if (!wr(out + "/counts.bin", counts.data(), 4 * ((size_t)n_defg + n_rxg))) return 3;

    std::vector<uint32_t> map(N);
    if ((rc = score_gm_get_grainmap(h, map.data(), SCORE_GM_MAP_RAW2GID))) return die("score_gm_get_grainmap", rc);
    if (!wr(out + "/map.bin", map.data(), 4 * N)) return 3;

    size_t nt = 0;
    if ((rc = score_gm_get_trajectory(h, NULL, 0, &nt))) return die("score_gm_get_trajectory", rc);
    std::vector<score_gm_traj_row> traj(nt + 1);
    if ((rc = score_gm_get_trajectory(h, traj.data(), nt, &nt))) return die("score_gm_get_trajectory", rc);
    if (!wr(out + "/traj.bin", traj.data(), nt * sizeof(score_gm_traj_row))) return 3;

    // mygrainevolution
    size_t ng = 0;
    if ((rc = score_gm_get_grainlog(h, NULL, NULL, 0, &ng))) return die("score_gm_get_grainlog", rc);
    std::vector<score_gm_grainlog_row> glog(ng + 1);
    std::vector<uint32_t> gcounts((ng + 1) * ((size_t)n_defg + n_rxg));
    if ((rc = score_gm_get_grainlog(h, glog.data(), gcounts.data(), ng, &ng))) return die("score_gm_get_grainlog", rc);
    if (!wr(out + "/glog.bin", glog.data(), ng * sizeof(score_gm_grainlog_row)) ||
        !wr(out + "/glogcounts.bin", gcounts.data(), 4 * ng * ((size_t)n_defg + n_rxg)))
        return 3;

    // cellStatii after the run (ompcrit_clarify_status_infectedcells)
    uint32_t ns = 0;
    if ((rc = score_gm_get_status(h, 0.5, 0, &ns, NULL, NULL))) return die("score_gm_get_status", rc);
    std::vector<uint32_t> st(2 * (size_t)ns + 2);
    if ((rc = score_gm_get_status(h, 0.5, ns, &ns, st.data(), st.data() + ns))) return die("score_gm_get_status", rc);
    if (!wr(out + "/status.bin", st.data(), 8 * (size_t)ns)) return 3;
    score_gm_destroy(h);

    // the final structure in SCORE's files (SCORE_Kernel.cpp:6755-6771; raw as
    // ompcrit_write_voxeldata_coloring_grainids names it)
    char name[512];
    if (score_io_raw_name(name, sizeof name, 1, 0, res.step, res.X, "GROWTH.3DGrainID") != SCORE_IO_OK) return 4;
    if (score_io_write_raw((out + "/" + name).c_str(), map.data(), nx, ny, nz) != SCORE_IO_OK) return 4;
    char h5[256], xmf[256];
    if (score_io_h5_name(h5, sizeof h5, 1) != SCORE_IO_OK || score_io_xdmf_name(xmf, sizeof xmf, 1) != SCORE_IO_OK) return 4;
    if (score_io_write_h5_grainid((out + "/" + h5).c_str(), "/GROWTH3DGrainID", res.step, map.data(), nx, ny, nz,
                                  NULL) != SCORE_IO_OK)
        return 4;
    const uint32_t snap[4] = {nx, ny, nz, res.step};
    if (score_io_write_xdmf((out + "/" + xmf).c_str(), h5, "/GROWTH3DGrainID", snap, 1) != SCORE_IO_OK) return 4;
    printf("caller ok: %u steps, X=%.6f, %zu grain logs, %u status entries\n", res.step, res.X, ng, ns);
    return 0;
}
