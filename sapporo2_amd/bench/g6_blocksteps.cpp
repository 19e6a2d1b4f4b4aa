// g6_blocksteps.cpp -- block-step benchmark driver over the GRAPE-6 C-ABI.
//
// A GRAPE-6 caller in the manner of the reference's own performance driver
// (testCodes/src/test_performance_blockStep.cpp:160-183): per block step it
// re-sends the previous block's particles with g6_set_j_particle_, sets the
// block time with g6_set_ti_ and evaluates the active particles in calls of
// at most g6_npipes_() i-particles: g6calc_firsthalf_, g6calc_lasthalf2_ and,
// with neighbour radii, g6_read_neighbour_list_ followed by
// g6_get_neighbour_list_ for every i-particle.  Every timed interval is taken
// with std::chrono around the C-ABI calls; nothing else runs in between.
//
// usage: g6_blocksteps <workload.bin> <steps> <warmup> <results.bin> [sample_block] [timing]
//   workload.bin  written by bench.py (layout in read_workload below)
//   results.bin   the sample block's results of the last timed step and the
//                 predicted j-state, read back by bench.py for the accuracy checks
//   sample_block  block index whose results are kept (default: the last block)
//   timing        1 (default): the engine's HIP events report each call's device
//                 window and kernel times; 0: off; 2: interleaved, 2 x steps timed
//                 steps alternating off (even) and on (odd), so that the API rate
//                 (from the uninstrumented steps) and the device times (from the
//                 instrumented ones) come from the same process and GPU state
// Prints one JSON line on stdout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "g6lib.h"
#include "sapporo_amd.h"

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Workload {
  int n = 0, nblocks = 0, resend = 0, ngb = 0;
  double eps2 = 0, tblock = 0;
  std::vector<int> id;
  std::vector<double> m, x, v, a, j, tj, h2, px, pv;  // vectors are n x 3
  std::vector<std::vector<int>> blocks;               // active particle indices per block
};

// Layout (little-endian): char magic[8] = "SAPG6WL1"; int32 n, nblocks, resend,
// ngb; double eps2, tblock; int32 id[n] (padded to 8 bytes); double m[n],
// x[3n], v[3n], a[3n], j[3n], tj[n], h2[n], px[3n], pv[3n]; int32
// count[nblocks], then the blocks' indices concatenated.
bool read_workload(const char* path, Workload& w) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  char magic[8];
  int32_t hdr[4];
  double hd[2];
  bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "SAPG6WL1", 8) && fread(hdr, 4, 4, f) == 4 &&
            fread(hd, 8, 2, f) == 2;
  if (ok) {
    w.n = hdr[0]; w.nblocks = hdr[1]; w.resend = hdr[2]; w.ngb = hdr[3];
    w.eps2 = hd[0]; w.tblock = hd[1];
    ok = w.n > 0 && w.nblocks > 0;
  }
  auto rd = [&](void* p, size_t sz, size_t cnt) { ok = ok && fread(p, sz, cnt, f) == cnt; };
  if (ok) {
    const size_t n = w.n;
    w.id.resize(n);
    rd(w.id.data(), 4, n);
    if (n % 2) { int32_t pad; rd(&pad, 4, 1); }
    for (auto* vec : {&w.m, &w.tj, &w.h2}) vec->resize(n);
    for (auto* vec : {&w.x, &w.v, &w.a, &w.j, &w.px, &w.pv}) vec->resize(3 * n);
    rd(w.m.data(), 8, n); rd(w.x.data(), 8, 3 * n); rd(w.v.data(), 8, 3 * n); rd(w.a.data(), 8, 3 * n);
    rd(w.j.data(), 8, 3 * n); rd(w.tj.data(), 8, n); rd(w.h2.data(), 8, n); rd(w.px.data(), 8, 3 * n);
    rd(w.pv.data(), 8, 3 * n);
    std::vector<int32_t> cnt(w.nblocks);
    rd(cnt.data(), 4, w.nblocks);
    w.blocks.resize(w.nblocks);
    for (int b = 0; ok && b < w.nblocks; ++b) {
      ok = cnt[b] >= 0 && cnt[b] <= w.n;
      if (ok) {
        w.blocks[b].resize(cnt[b]);
        rd(w.blocks[b].data(), 4, cnt[b]);
      }
      for (int i : w.blocks[b]) ok = ok && i >= 0 && i < w.n;
    }
  }
  fclose(f);
  return ok;
}

// The caller's contiguous i-arrays of one block (an integrator predicts its
// active particles into such arrays before the force call).
struct BlockIn {
  std::vector<int> ids;
  std::vector<double> xi, vi, h2;
};

// Per-call measurements.
struct CallRec {
  int step, block, ni, P;
  bool inst;  // the step ran with the engine's HIP events on
  double first_s, wait_s, read_s, copy_s;  // firsthalf, lasthalf2, read_neighbour_list, get_neighbour_list loop
  float dev_ms, force_ms, pred_ms, scat_ms, fin_ms;
  int nupd;
};

double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t k = v.size() / 2;
  return v.size() % 2 ? v[k] : 0.5 * (v[k - 1] + v[k]);
}

void send_j(const Workload& w, int i, double tj, const double* x, const double* v) {
  int cl = 0, addr = i, index = w.id[i];
  double dtj = 0.0, mass = w.m[i];
  double k18[3] = {0, 0, 0};
  double a2[3], j6[3], xx[3], vv[3];
  for (int d = 0; d < 3; ++d) {
    a2[d] = 0.5 * w.a[3 * i + d];
    j6[d] = w.j[3 * i + d] / 6.0;
    xx[d] = x[3 * i + d];
    vv[d] = v[3 * i + d];
  }
  g6_set_j_particle_(&cl, &addr, &index, &tj, &dtj, &mass, k18, j6, a2, vv, xx);
}

void json_array(FILE* o, const char* name, const std::vector<double>& v, bool comma = true) {
  fprintf(o, "\"%s\": [", name);
  for (size_t k = 0; k < v.size(); ++k) fprintf(o, "%s%.9g", k ? ", " : "", v[k]);
  fprintf(o, "]%s", comma ? ", " : "");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s <workload.bin> <steps> <warmup> <results.bin> [sample_block] [timing]\n", argv[0]);
    return 2;
  }
  Workload w;
  if (!read_workload(argv[1], w)) {
    fprintf(stderr, "g6_blocksteps: cannot read workload %s\n", argv[1]);
    return 2;
  }
  const int steps = std::max(1, atoi(argv[2]));
  const int warmup = std::max(0, atoi(argv[3]));
  const int sample_block = argc > 5 ? atoi(argv[5]) : w.nblocks - 1;
  const int timing = argc > 6 ? atoi(argv[6]) : 1;
  if (sample_block < 0 || sample_block >= w.nblocks) {
    fprintf(stderr, "g6_blocksteps: sample block out of range\n");
    return 2;
  }

  int cl = 0;
  if (g6_open_(&cl) != 0) {
    fprintf(stderr, "g6_blocksteps: g6_open_ failed\n");
    return 1;
  }
  sapamd_engine* h = sapamd_g6_engine();
  sapamd_set_timing(h, timing == 1);
  const int total = timing == 2 ? 2 * steps : steps;  // timed steps
  const int npipes = g6_npipes_();
  int nj = w.n;

  // caller-side arrays: per block, the active particles' ids, predicted
  // positions / velocities and neighbour radii (zero without lists)
  std::vector<BlockIn> in(w.nblocks);
  for (int b = 0; b < w.nblocks; ++b) {
    const auto& sel = w.blocks[b];
    BlockIn& bi = in[b];
    bi.ids.resize(sel.size());
    bi.xi.resize(3 * sel.size());
    bi.vi.resize(3 * sel.size());
    bi.h2.resize(sel.size());
    for (size_t k = 0; k < sel.size(); ++k) {
      const int i = sel[k];
      bi.ids[k] = w.id[i];
      for (int d = 0; d < 3; ++d) {
        bi.xi[3 * k + d] = w.px[3 * i + d];
        bi.vi[3 * k + d] = w.pv[3 * i + d];
      }
      bi.h2[k] = w.ngb ? w.h2[i] : 0.0;
    }
  }
  size_t maxb = 0;
  for (auto& sel : w.blocks) maxb = std::max(maxb, sel.size());
  std::vector<double> acc(3 * maxb), jrk(3 * maxb), pot(maxb);
  std::vector<int> inn(maxb);
  std::vector<int> lbuf(256);
  int maxlen = 256;

  // initial j-particles: raw state with their own times
  for (int i = 0; i < w.n; ++i) send_j(w, i, w.tj[i], w.x.data(), w.v.data());
  fprintf(stderr, "g6_blocksteps: %d j-particles sent; %d + %d steps of %d blocks\n", w.n, warmup, steps,
          w.nblocks);

  // kept sample: results of sample_block in the last timed step
  const size_t ns = w.blocks[sample_block].size();
  std::vector<double> s_acc(3 * ns), s_jrk(3 * ns), s_pot(ns);
  std::vector<int> s_inn(ns), s_cnt(ns);
  std::vector<int> s_list;

  std::vector<CallRec> recs;
  std::vector<double> step_api, step_resend, step_copy, step_setti;
  std::vector<char> step_inst;
  const Clock::time_point run0 = Clock::now();
  for (int step = 0; step < warmup + total; ++step) {
    const bool timed = step >= warmup;
    const bool keep = step == warmup + total - 1;
    const bool inst = timing == 1 || (timing == 2 && step % 2 == 1);
    if (timing == 2) sapamd_set_timing(h, inst ? 1 : 0);
    double api = 0, resend = 0, copy = 0, setti = 0;
    for (int b = 0; b < w.nblocks; ++b) {
      const auto& sel = w.blocks[b];
      if (w.resend) {
        const auto& prev = w.blocks[(b + w.nblocks - 1) % w.nblocks];
        const Clock::time_point r0 = Clock::now();
        for (int i : prev) send_j(w, i, w.tblock, w.px.data(), w.pv.data());
        resend += secs(r0, Clock::now());
      }
      const Clock::time_point s0 = Clock::now();
      double tb = w.tblock;
      g6_set_ti_(&cl, &tb);
      const double st = secs(s0, Clock::now());
      setti += st;
      api += st;
      BlockIn& bi = in[b];
      for (size_t c0 = 0; c0 < sel.size(); c0 += npipes) {
        int ni = (int)std::min<size_t>(npipes, sel.size() - c0);
        double eps2 = w.eps2;
        auto X = reinterpret_cast<double(*)[3]>(&bi.xi[3 * c0]);
        auto V = reinterpret_cast<double(*)[3]>(&bi.vi[3 * c0]);
        auto A = reinterpret_cast<double(*)[3]>(&acc[3 * c0]);
        auto J = reinterpret_cast<double(*)[3]>(&jrk[3 * c0]);
        int* ids = &bi.ids[c0];
        double* h2 = &bi.h2[c0];
        const Clock::time_point t0 = Clock::now();
        g6calc_firsthalf_(&cl, &nj, &ni, ids, X, V, A, J, &pot[c0], &eps2, h2);
        const Clock::time_point t1 = Clock::now();
        g6calc_lasthalf2_(&cl, &nj, &ni, ids, X, V, &eps2, h2, A, J, &pot[c0], &inn[c0]);
        const Clock::time_point t2 = Clock::now();
        if (w.ngb) g6_read_neighbour_list_(&cl);
        const Clock::time_point t3 = Clock::now();
        if (w.ngb) {
          for (int i = 0; i < ni; ++i) {
            int ip = i, nn = 0;
            g6_get_neighbour_list_(&cl, &ip, &maxlen, &nn, lbuf.data());
            if (keep && b == sample_block) {
              s_cnt[c0 + i] = nn;
              s_list.insert(s_list.end(), lbuf.begin(), lbuf.begin() + std::min(nn, maxlen));
            }
          }
        }
        const Clock::time_point t4 = Clock::now();
        api += secs(t0, t3);
        copy += secs(t3, t4);
        if (timed) {
          CallRec r{};
          r.step = step - warmup; r.block = b; r.ni = ni; r.inst = inst;
          r.first_s = secs(t0, t1); r.wait_s = secs(t1, t2); r.read_s = secs(t2, t3); r.copy_s = secs(t3, t4);
          if (inst) {
            r.dev_ms = sapamd_last_device_ms(h);
            r.force_ms = sapamd_last_force_ms(h);
            r.pred_ms = sapamd_last_predict_ms(h);
            r.scat_ms = sapamd_last_scatter_ms(h, &r.nupd);
            r.fin_ms = sapamd_last_finalize_ms(h);
          }
          sapamd_last_config(h, &r.P, nullptr, nullptr);
          recs.push_back(r);
        }
      }
      if (keep && b == sample_block) {
        std::copy(acc.begin(), acc.begin() + 3 * ns, s_acc.begin());
        std::copy(jrk.begin(), jrk.begin() + 3 * ns, s_jrk.begin());
        std::copy(pot.begin(), pot.begin() + ns, s_pot.begin());
        std::copy(inn.begin(), inn.begin() + ns, s_inn.begin());
      }
    }
    if (timed) {
      step_api.push_back(api);
      step_resend.push_back(resend);
      step_copy.push_back(copy);
      step_setti.push_back(setti);
      step_inst.push_back(inst ? 1 : 0);
    }
  }
  const double run_s = secs(run0, Clock::now());
  fprintf(stderr, "g6_blocksteps: timed loop done (%.3f s)\n", run_s);

  // predicted j-state the force calls used (raw and predicted read-back)
  std::vector<double> ppos(3 * (size_t)w.n), pvel(3 * (size_t)w.n);
  for (int i = 0; i < w.n; ++i) {
    double pos[3], vel[3], a[3], jk[3], mass, e2;
    int idv;
    get_j_part_data(i, w.n, pos, vel, a, jk, &ppos[3 * (size_t)i], &pvel[3 * (size_t)i], mass, e2, idv);
  }
  g6_close_(&cl);
  fprintf(stderr, "g6_blocksteps: read-back done, library closed\n");

  // results file
  {
    FILE* f = fopen(argv[4], "wb");
    if (!f) {
      fprintf(stderr, "g6_blocksteps: cannot write %s\n", argv[4]);
      return 1;
    }
    const int32_t hdr[4] = {(int32_t)ns, (int32_t)s_list.size(), w.n, w.ngb};
    fwrite("SAPG6RS1", 1, 8, f);
    fwrite(hdr, 4, 4, f);
    fwrite(s_acc.data(), 8, s_acc.size(), f);
    fwrite(s_jrk.data(), 8, s_jrk.size(), f);
    fwrite(s_pot.data(), 8, s_pot.size(), f);
    fwrite(s_inn.data(), 4, s_inn.size(), f);
    fwrite(s_cnt.data(), 4, s_cnt.size(), f);
    fwrite(s_list.data(), 4, s_list.size(), f);
    fwrite(ppos.data(), 8, ppos.size(), f);
    fwrite(pvel.data(), 8, pvel.size(), f);
    fclose(f);
  }

  // summary: per step, and per block (medians over the timed steps).  With
  // interleaved timing the API figures come from the uninstrumented steps and
  // the device and kernel times from the instrumented ones.
  double inter_step = 0;
  for (auto& sel : w.blocks) inter_step += (double)sel.size() * w.n;
  auto of_kind = [&](int s, bool inst) { return timing != 2 || (step_inst[s] != 0) == inst; };
  auto pick = [&](const std::vector<double>& v, bool inst) {
    std::vector<double> out;
    for (int s = 0; s < (int)v.size(); ++s)
      if (of_kind(s, inst)) out.push_back(v[s]);
    return out;
  };
  std::vector<double> step_dev(total, 0.0), step_f64(total, 0.0), step_first(total, 0.0), step_wait(total, 0.0),
      step_read(total, 0.0);
  double f64_inter = 0, f64_ms = 0, pred_ms = 0, scat_ms = 0, scat_upd = 0;
  long f64_launches = 0, pred_launches = 0, scat_launches = 0;
  for (const CallRec& r : recs) {
    step_first[r.step] += r.first_s;
    step_wait[r.step] += r.wait_s;
    step_read[r.step] += r.read_s;
    if (!r.inst) continue;
    if (r.pred_ms > 0) {
      pred_ms += r.pred_ms;
      ++pred_launches;
    }
    if (r.nupd > 0) {
      scat_ms += r.scat_ms;
      scat_upd += r.nupd;
      ++scat_launches;
    }
    step_dev[r.step] += r.dev_ms;
    if (r.P == 64) {
      step_f64[r.step] += r.force_ms;
      f64_inter += (double)r.ni * w.n;
      f64_ms += r.force_ms;
      ++f64_launches;
    }
  }
  FILE* o = stdout;
  fprintf(o, "{\"driver\": \"g6_blocksteps\", \"n\": %d, \"nblocks\": %d, \"steps\": %d, \"warmup\": %d, ", w.n,
          w.nblocks, steps, warmup);
  fprintf(o, "\"timing\": %d, \"interactions_per_step\": %.17g, \"run_s\": %.6f, ", timing, inter_step, run_s);
  fprintf(o, "\"api_median_s\": %.9g, ", median(pick(step_api, false)));
  json_array(o, "step_api_s", pick(step_api, false));
  json_array(o, "step_api_instrumented_s", pick(step_api, true));
  json_array(o, "step_setti_s", pick(step_setti, false));
  json_array(o, "step_resend_s", pick(step_resend, false));
  json_array(o, "step_list_copy_s", pick(step_copy, false));
  json_array(o, "step_firsthalf_s", pick(step_first, false));
  json_array(o, "step_lasthalf_s", pick(step_wait, false));
  json_array(o, "step_read_ngb_s", pick(step_read, false));
  json_array(o, "step_device_ms", pick(step_dev, true));
  json_array(o, "step_force64_ms", pick(step_f64, true));
  fprintf(o, "\"force64\": {\"launches\": %ld, \"interactions\": %.17g, \"ms\": %.9g}, ", f64_launches, f64_inter,
          f64_ms);
  fprintf(o, "\"predict\": {\"launches\": %ld, \"ms\": %.9g, \"nj\": %d}, ", pred_launches, pred_ms, w.n);
  fprintf(o, "\"scatter\": {\"launches\": %ld, \"ms\": %.9g, \"updates\": %.17g}, ", scat_launches, scat_ms,
          scat_upd);
  fprintf(o, "\"blocks\": [");
  for (int b = 0; b < w.nblocks; ++b) {
    std::vector<double> api(total, 0.0), dev(total, 0.0), fst(total, 0.0), wt(total, 0.0), rd(total, 0.0),
        fo(total, 0.0), fz(total, 0.0);
    int calls = 0, P = 0;
    for (const CallRec& r : recs) {
      if (r.block != b) continue;
      api[r.step] += r.first_s + r.wait_s + r.read_s;
      dev[r.step] += r.dev_ms;
      fo[r.step] += r.force_ms;
      fz[r.step] += r.fin_ms;
      fst[r.step] += r.first_s;
      wt[r.step] += r.wait_s;
      rd[r.step] += r.read_s;
      if (r.step == 0) ++calls;
      P = r.P;
    }
    fprintf(o,
            "%s{\"ni\": %zu, \"calls\": %d, \"P\": %d, \"api_ms\": %.6g, \"device_ms\": %.6g, \"force_ms\": %.6g, "
            "\"finalize_ms\": %.6g, \"firsthalf_ms\": %.6g, \"lasthalf_ms\": %.6g, \"read_ngb_ms\": %.6g}",
            b ? ", " : "", w.blocks[b].size(), calls, P, 1e3 * median(pick(api, false)), median(pick(dev, true)),
            median(pick(fo, true)), median(pick(fz, true)), 1e3 * median(pick(fst, false)),
            1e3 * median(pick(wt, false)), 1e3 * median(pick(rd, false)));
  }
  fprintf(o, "]}\n");
  fflush(o);
  fprintf(stderr, "g6_blocksteps: done\n");
  return 0;
}
