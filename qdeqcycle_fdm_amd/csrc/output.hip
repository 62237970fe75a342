// output.hip -- the reference's fault output records and an event catalogue.
//
// Records (odeSolver.cpp:1442-1457, 1638-1652): output_fault_0.bin holds
// t, V, Slip, Ts, psi (nz each), U (nx*nz), Tn, a, b (nz each), dt after the
// t = 0 stage; output_fault_<k>.bin, written after every `every`-th accepted
// step (it_save_jump, SetParams.cpp:27), holds t, V, Slip, Ts, psi, dt.  All
// FP64, native byte order, no headers.  Records after the first are copied
// from the device into a ring of pinned host slots on the context's stream
// and written to disk when their copy has completed, so stepping never waits
// on the file system.
//
// Events (no reference counterpart; the reference only prints V_max per
// step, odeSolver.cpp:1636): after every accepted step one small kernel
// updates a device-side catalogue: an event starts at the first step whose
// max slip rate exceeds the threshold and ends at the first step back at or
// below it; each entry keeps the onset time and step, the node of max V at
// onset, the peak V and its time and node, the end, and the largest slip
// accumulated on any node during the event.

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

namespace {

constexpr int EV_THREADS = 1024;

struct DevEvent {
  double t_onset, t_end, t_peak, v_peak, slip_max;
  int it_onset, it_end, node_onset, node_peak;
};

// state[0] = in_event, state[1] = events recorded (incl. the open one)
__global__ __launch_bounds__(EV_THREADS) void k_events(
    const double *__restrict__ V, const double *__restrict__ slip, int nk,
    double t, int it, double vthr, int *__restrict__ state,
    DevEvent *__restrict__ ev, int cap, double *__restrict__ slip0) {
  __shared__ double sv[EV_THREADS];
  __shared__ int si[EV_THREADS];
  double best = -1.0;
  int bi = 1;
  for (int k = 1 + (int)threadIdx.x; k <= nk; k += EV_THREADS) {
    const double v = V[k];
    if (v > best) {
      best = v;
      bi = k;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = EV_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = sv[threadIdx.x + s];
      const int oi = si[threadIdx.x + s];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const double vmax = sv[0];
  const int imax = si[0];
  const int in_event = state[0];
  const int n = state[1];
  __syncthreads();
  if (!in_event && vmax > vthr) {
    // onset: remember the slip, open an entry
    for (int k = threadIdx.x; k <= nk + 1; k += EV_THREADS) slip0[k] = slip[k];
    if (threadIdx.x == 0) {
      if (n < cap) {
        DevEvent e;
        e.t_onset = t;
        e.t_end = -1.0;
        e.t_peak = t;
        e.v_peak = vmax;
        e.slip_max = 0.0;
        e.it_onset = it;
        e.it_end = -1;
        e.node_onset = imax;
        e.node_peak = imax;
        ev[n] = e;
      }
      state[0] = 1;
      state[1] = n + 1;
    }
    return;
  }
  if (!in_event) return;
  // inside an event: peak and accumulated slip
  double ds = 0.0;
  for (int k = 1 + (int)threadIdx.x; k <= nk; k += EV_THREADS)
    ds = fmax(ds, slip[k] - slip0[k]);
  sv[threadIdx.x] = ds;
  __syncthreads();
  for (int s = EV_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sv[threadIdx.x] = fmax(sv[threadIdx.x], sv[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && n >= 1 && n <= cap) {
    DevEvent &e = ev[n - 1];
    e.slip_max = fmax(e.slip_max, sv[0]);
    if (vmax > e.v_peak) {
      e.v_peak = vmax;
      e.t_peak = t;
      e.node_peak = imax;
    }
    if (!(vmax > vthr)) {
      e.t_end = t;
      e.it_end = it;
    }
  }
  if (threadIdx.x == 0 && !(vmax > vthr)) state[0] = 0;
}

int write_file(const std::string &path, const double *data, size_t n,
               std::string *err) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) {
    *err = "cannot open " + path;
    return QDEQ_ERR_ARG;
  }
  const size_t w = std::fwrite(data, sizeof(double), n, fp);
  const int rc = std::fclose(fp);
  if (w != n || rc != 0) {
    *err = "short write to " + path;
    return QDEQ_ERR_ARG;
  }
  return QDEQ_OK;
}

}  // namespace

struct OutputRing {
  std::string dir;
  int every = 0;
  int next_index = 1;        // it_save of the next periodic record
  size_t slot_len = 0;       // doubles per record: 4*nz + 2
  std::vector<double *> slot;
  std::vector<hipEvent_t> done;
  std::vector<int> index;    // record index held by a slot, -1 if free
  std::vector<int> fifo;     // slots in write order
};

// Writes every pending record whose copy has completed (all of them when
// block), oldest first.
static int output_drain(qdeq_ctx *c, bool block) {
  OutputRing *o = c->out;
  while (!o->fifo.empty()) {
    const int s = o->fifo.front();
    if (block) {
      QDEQ_HIP(c, hipEventSynchronize(o->done[(size_t)s]));
    } else {
      const hipError_t e = hipEventQuery(o->done[(size_t)s]);
      if (e == hipErrorNotReady) break;
      QDEQ_HIP(c, e);
    }
    std::string err;
    const std::string path =
        o->dir + "/output_fault_" + std::to_string(o->index[(size_t)s]) + ".bin";
    const int st = write_file(path, o->slot[(size_t)s], o->slot_len, &err);
    if (st != QDEQ_OK) return set_error(c, st, err);
    o->index[(size_t)s] = -1;
    o->fifo.erase(o->fifo.begin());
  }
  return QDEQ_OK;
}

// Enqueues the record of the current state (after an accepted step).
static int output_record(qdeq_ctx *c) {
  OutputRing *o = c->out;
  int st = output_drain(c, false);
  if (st != QDEQ_OK) return st;
  int s = -1;
  for (size_t i = 0; i < o->slot.size(); i++)
    if (o->index[i] < 0) {
      s = (int)i;
      break;
    }
  if (s < 0) {  // ring full: finish the oldest
    const int oldest = o->fifo.front();
    QDEQ_HIP(c, hipEventSynchronize(o->done[(size_t)oldest]));
    if ((st = output_drain(c, false)) != QDEQ_OK) return st;
    s = oldest;
  }
  const size_t nz = (size_t)c->nz;
  double *p = o->slot[(size_t)s];
  p[0] = c->t;
  p[1 + 4 * nz] = c->dt;
  const double *src[4] = {c->fb.FV[1], c->fb.slip, c->fb.Ts, c->fb.psi};
  for (int a = 0; a < 4; a++)
    QDEQ_HIP(c, hipMemcpyAsync(p + 1 + a * nz, src[a], nz * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipEventRecord(o->done[(size_t)s], c->stream));
  o->index[(size_t)s] = o->next_index++;
  o->fifo.push_back(s);
  return QDEQ_OK;
}

int output_after_step(qdeq_ctx *c) {
  if (!c->out || c->it % c->out->every != 0) return QDEQ_OK;
  return output_record(c);
}

void output_free(qdeq_ctx *c) {
  if (!c->out) return;
  for (double *p : c->out->slot) (void)hipHostFree(p);
  for (hipEvent_t e : c->out->done) (void)hipEventDestroy(e);
  delete c->out;
  c->out = nullptr;
}

void events_after_step(qdeq_ctx *c) {
  hipLaunchKernelGGL(k_events, dim3(1), dim3(EV_THREADS), 0, c->stream,
                     c->fb.FV[1], c->fb.slip, c->cfg.nk, c->t, c->it, c->ev_thr,
                     c->ev_state, reinterpret_cast<DevEvent *>(c->ev_table),
                     c->ev_cap, c->ev_slip0);
}

void events_free(qdeq_ctx *c) {
  (void)hipFree(c->ev_state);
  (void)hipFree(c->ev_table);
  (void)hipFree(c->ev_slip0);
  c->ev_state = nullptr;
  c->ev_table = nullptr;
  c->ev_slip0 = nullptr;
  c->ev_on = false;
}

}  // namespace qdeq

using namespace qdeq;

static_assert(sizeof(qdeq_event) == sizeof(qdeq::DevEvent),
              "qdeq_event must match the device layout");

extern "C" {

int qdeq_output_open(qdeq_ctx *c, const char *dir, int every) {
  if (!c || !dir || every < 1) return set_error(c, QDEQ_ERR_ARG, "bad argument");
  int st;
  if (!c->rk_ready) {
    // record 0 follows the t = 0 stage (odeSolver.cpp:1428-1457)
    if ((st = qdeq_rk43_step(c, 0, nullptr)) != QDEQ_OK) return st;
  }
  if (c->out && (st = qdeq_output_close(c)) != QDEQ_OK) return st;
  const size_t nz = (size_t)c->nz, ng = (size_t)c->nx * c->nz;
  // record 0: t, V, Slip, Ts, psi, U, Tn, a, b, dt
  std::vector<double> rec(1 + 4 * nz + ng + 3 * nz + 1);
  rec[0] = c->t;
  double *V = rec.data() + 1, *slip = V + nz, *Ts = slip + nz, *psi = Ts + nz;
  double *U = psi + nz, *Tn = U + ng, *a = Tn + nz, *b = a + nz;
  if ((st = qdeq_rk43_get_state(c, nullptr, U, slip, psi, V, nullptr, Ts))) return st;
  QDEQ_HIP(c, hipMemcpy(Tn, c->fb.Tn, nz * sizeof(double), hipMemcpyDeviceToHost));
  QDEQ_HIP(c, hipMemcpy(a, c->fb.a, nz * sizeof(double), hipMemcpyDeviceToHost));
  QDEQ_HIP(c, hipMemcpy(b, c->fb.b, nz * sizeof(double), hipMemcpyDeviceToHost));
  rec.back() = c->dt;
  std::string err;
  if ((st = write_file(std::string(dir) + "/output_fault_0.bin", rec.data(), rec.size(),
                       &err)) != QDEQ_OK)
    return set_error(c, st, err);
  OutputRing *o = new OutputRing();
  o->dir = dir;
  o->every = every;
  o->slot_len = 4 * nz + 2;
  const int nslots = 32;
  c->out = o;
  for (int i = 0; i < nslots; i++) {
    double *p = nullptr;
    hipEvent_t e = nullptr;
    if (hipHostMalloc(&p, o->slot_len * sizeof(double)) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      if (p) (void)hipHostFree(p);
      output_free(c);
      return set_error(c, QDEQ_ERR_NOMEM, "output ring allocation");
    }
    o->slot.push_back(p);
    o->done.push_back(e);
    o->index.push_back(-1);
  }
  return QDEQ_OK;
}

int qdeq_output_close(qdeq_ctx *c) {
  if (!c) return QDEQ_ERR_ARG;
  if (!c->out) return QDEQ_OK;
  const int st = output_drain(c, true);
  output_free(c);
  return st;
}

int qdeq_events_enable(qdeq_ctx *c, double v_threshold, int capacity) {
  if (!c || !(v_threshold > 0) || capacity < 1)
    return set_error(c, QDEQ_ERR_ARG, "bad argument");
  events_free(c);
  const size_t nz = (size_t)c->nz;
  QDEQ_HIP(c, hipMalloc(&c->ev_state, 2 * sizeof(int)));
  QDEQ_HIP(c, hipMalloc(&c->ev_table, (size_t)capacity * sizeof(DevEvent)));
  QDEQ_HIP(c, hipMalloc(&c->ev_slip0, nz * sizeof(double)));
  QDEQ_HIP(c, hipMemsetAsync(c->ev_state, 0, 2 * sizeof(int), c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  c->ev_thr = v_threshold;
  c->ev_cap = capacity;
  c->ev_on = true;
  return QDEQ_OK;
}

int qdeq_events_get(qdeq_ctx *c, qdeq_event *out, int max, int *n) {
  if (!c || !n || (max > 0 && !out)) return set_error(c, QDEQ_ERR_ARG, "bad argument");
  if (!c->ev_on) return set_error(c, QDEQ_ERR_STATE, "events not enabled");
  int state[2] = {0, 0};
  QDEQ_HIP(c, hipMemcpyAsync(state, c->ev_state, sizeof(state), hipMemcpyDeviceToHost,
                             c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  const int stored = state[1] < c->ev_cap ? state[1] : c->ev_cap;
  *n = state[1];
  const int m = stored < max ? stored : max;
  if (m > 0)
    QDEQ_HIP(c, hipMemcpy(out, c->ev_table, (size_t)m * sizeof(DevEvent),
                          hipMemcpyDeviceToHost));
  return QDEQ_OK;
}

}  // extern "C"
