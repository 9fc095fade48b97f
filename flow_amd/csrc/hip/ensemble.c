/*
 * ensemble.c -- host side of the batching form (flow_hip.h).
 *
 * An ensemble holds many independent meshes of uneven sizes in HBM. Each
 * instance has the 16 arrays of solve_hydro_2d, its own geometry (arch's
 * uniform spacing, main.c:20-31 with the instance's sizes), its own dt pair
 * and its own totals. One step of the whole ensemble is seven launches
 * (fk_ens_step), optionally replayed from a captured hipGraph per sweep
 * parity; per-step dt pairs and totals go to device history buffers and
 * reach the host once per run.
 *
 * The autotuner times short runs of each candidate chunk height and work
 * order from a snapshot of the state, keeps the fastest, and restores the
 * state.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include "../../../include/flow_interface.h"
#include "flowhip_ctx.h"

#define ENS_ORDER_INSTANCE 0 /* items grouped by instance, in instance order */
#define ENS_ORDER_LARGEST 1  /* instances by decreasing size */
#define ENS_ORDER_ROUND 2    /* one chunk row of every instance in turn */

struct FlowHipEnsemble {
  int ninst, fp32, esize;
  int* gnx;
  int* gny;
  FkEnsInst* h_insts;
  FkEnsInst* d_insts;
  void* pool;          /* all arrays of all instances */
  size_t pool_bytes;
  void* snap;          /* autotuner snapshot of pool + dt pairs */
  void* geo;           /* geometry of all instances */
  double* d_dtp;       /* 2 per instance */
  double* h_dtp;       /* pinned */
  int* d_work;
  int nwork;
  int chunk_rows, order;
  int pow2; /* every instance's spacings are powers of two */
  int uniform; /* spacings read once per work item */
  int fused_momentum;
  unsigned int* d_queue; /* work-queue counters of the ensemble kernels */
  FkEnsPartials part;
  int part_items;
  FkEnsHistory hist;
  int hist_rows;
  hipStream_t stream;
  int use_graphs;
  hipGraphExec_t graph[2];
  int graph_ok[2];
};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static void ens_drop_graphs(FlowHipEnsemble* E) {
  for (int k = 0; k < 2; ++k) {
    if (E->graph_ok[k]) HIPCHECK(hipGraphExecDestroy(E->graph[k]));
    E->graph_ok[k] = 0;
  }
}

/* Items per instance for the current chunk height, the partial buffers,
 * and the work list in the current order. */
static void ens_build_work(FlowHipEnsemble* E) {
  const int W = fk_ens_strip_width(), H = E->chunk_rows, n = E->ninst;
  int total = 0;
  for (int i = 0; i < n; ++i) {
    FkEnsInst* in = &E->h_insts[i];
    in->nstrips = (E->gnx[i] + W - 1) / W;
    in->nchunks = (E->gny[i] + H - 1) / H;
    in->nitems = in->nstrips * in->nchunks;
    in->item0 = total;
    total += in->nitems;
  }
  int* pairs = (int*)malloc(sizeof(int) * 2 * (size_t)total);
  int* idx = (int*)malloc(sizeof(int) * (size_t)n);
  if (!pairs || !idx) TERMINATE("ensemble: out of host memory");
  for (int i = 0; i < n; ++i) idx[i] = i;
  if (E->order == ENS_ORDER_LARGEST) {
    /* insertion sort by decreasing item count (stable) */
    for (int a = 1; a < n; ++a) {
      const int v = idx[a];
      int b = a - 1;
      while (b >= 0 && E->h_insts[idx[b]].nitems < E->h_insts[v].nitems) {
        idx[b + 1] = idx[b];
        --b;
      }
      idx[b + 1] = v;
    }
  }
  int w = 0;
  if (E->order == ENS_ORDER_ROUND) {
    int maxc = 0;
    for (int i = 0; i < n; ++i)
      if (E->h_insts[i].nchunks > maxc) maxc = E->h_insts[i].nchunks;
    for (int c = 0; c < maxc; ++c)
      for (int i = 0; i < n; ++i) {
        const FkEnsInst* in = &E->h_insts[i];
        if (c >= in->nchunks) continue;
        for (int s = 0; s < in->nstrips; ++s) {
          pairs[2 * w] = i;
          pairs[2 * w + 1] = c * in->nstrips + s;
          ++w;
        }
      }
  } else {
    for (int a = 0; a < n; ++a) {
      const int i = idx[a];
      for (int k = 0; k < E->h_insts[i].nitems; ++k) {
        pairs[2 * w] = i;
        pairs[2 * w + 1] = k;
        ++w;
      }
    }
  }
  E->nwork = total;
  HIPCHECK(hipFree(E->d_work));
  HIPCHECK(hipMalloc((void**)&E->d_work, sizeof(int) * 2 * (size_t)total));
  HIPCHECK(hipMemcpy(E->d_work, pairs, sizeof(int) * 2 * (size_t)total,
                     hipMemcpyHostToDevice));
  if (total > E->part_items) {
    HIPCHECK(hipFree(E->part.min));
    HIPCHECK(hipFree(E->part.nan));
    HIPCHECK(hipFree(E->part.sums));
    HIPCHECK(hipMalloc((void**)&E->part.min, sizeof(double) * (size_t)total));
    HIPCHECK(hipMalloc((void**)&E->part.nan, sizeof(int) * (size_t)total));
    HIPCHECK(hipMalloc((void**)&E->part.sums, sizeof(double) * 4 * (size_t)total));
    E->part_items = total;
  }
  HIPCHECK(hipMemcpy(E->d_insts, E->h_insts, sizeof(FkEnsInst) * (size_t)n,
                     hipMemcpyHostToDevice));
  free(pairs);
  free(idx);
  ens_drop_graphs(E);
}

static FkEnsLaunch ens_launch(const FlowHipEnsemble* E) {
  FkEnsLaunch L;
  L.insts = E->d_insts;
  L.work = E->d_work;
  L.ninst = E->ninst;
  L.nwork = E->nwork;
  L.chunk_rows = E->chunk_rows;
  L.pow2 = E->pow2;
  L.uniform = E->uniform;
  L.fused_momentum = E->fused_momentum;
  L.queue = E->d_queue;
  L.part = E->part;
  L.hist = E->hist;
  return L;
}

FlowHipEnsemble* flowhip_ensemble_new(int ninst, const int* global_nx,
                                      const int* global_ny, const double* width,
                                      const double* height, const double* max_dt,
                                      const double* dt_param, int fp32) {
  if (ninst < 1 || !global_nx || !global_ny || !width || !height || !max_dt ||
      !dt_param)
    return NULL;
  for (int i = 0; i < ninst; ++i)
    if (!flowhip_mesh_size_ok(global_nx[i], global_ny[i])) return NULL;
  FlowHipEnsemble* E = (FlowHipEnsemble*)calloc(1, sizeof(FlowHipEnsemble));
  if (!E) return NULL;
  E->ninst = ninst;
  E->fp32 = fp32 ? 1 : 0;
  E->esize = fp32 ? (int)sizeof(float) : (int)sizeof(double);
  E->chunk_rows = 48;
  E->order = ENS_ORDER_LARGEST;
  E->fused_momentum = 1;
  E->gnx = (int*)malloc(sizeof(int) * (size_t)ninst);
  E->gny = (int*)malloc(sizeof(int) * (size_t)ninst);
  E->h_insts = (FkEnsInst*)calloc((size_t)ninst, sizeof(FkEnsInst));
  if (!E->gnx || !E->gny || !E->h_insts) TERMINATE("ensemble: out of host memory");
  memcpy(E->gnx, global_nx, sizeof(int) * (size_t)ninst);
  memcpy(E->gny, global_ny, sizeof(int) * (size_t)ninst);
  HIPCHECK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));

  /* array pool: 16 arrays of (L+1)^2 per instance, 256-byte aligned */
  size_t off = 0, goff = 0;
  size_t* aoff = (size_t*)malloc(sizeof(size_t) * 16 * (size_t)ninst);
  size_t* gof = (size_t*)malloc(sizeof(size_t) * (size_t)ninst);
  if (!aoff || !gof) TERMINATE("ensemble: out of host memory");
  for (int i = 0; i < ninst; ++i) {
    const size_t lx = (size_t)global_nx[i] + 5, ly = (size_t)global_ny[i] + 5;
    for (int k = 0; k < 16; ++k) {
      aoff[16 * i + k] = off;
      off = align_up(off + lx * ly * (size_t)E->esize, 256);
    }
    gof[i] = goff;
    goff = align_up(goff + 3 * (lx + ly) * (size_t)E->esize, 256);
  }
  E->pool_bytes = off;
  HIPCHECK(hipMalloc(&E->pool, off));
  HIPCHECK(hipMemset(E->pool, 0, off));
  HIPCHECK(hipMalloc(&E->geo, goff));
  HIPCHECK(hipMalloc((void**)&E->d_dtp, sizeof(double) * 2 * (size_t)ninst));
  HIPCHECK(hipHostMalloc((void**)&E->h_dtp, sizeof(double) * 2 * (size_t)ninst,
                         hipHostMallocDefault));

  /* geometry: arch's uniform spacing (width / global_nx for every entry),
   * its reciprocals by a correctly rounded division, FP32 copies if asked */
  size_t maxl = 0;
  for (int i = 0; i < ninst; ++i) {
    const size_t l = (size_t)(global_nx[i] > global_ny[i] ? global_nx[i] : global_ny[i]) + 5;
    if (l > maxl) maxl = l;
  }
  double* hbuf = (double*)malloc(sizeof(double) * maxl);
  double* dtmp = NULL;
  HIPCHECK(hipMalloc((void**)&dtmp, sizeof(double) * 6 * maxl));
  if (!hbuf) TERMINATE("ensemble: out of host memory");
  for (int i = 0; i < ninst; ++i) {
    const int lx = global_nx[i] + 5, ly = global_ny[i] + 5;
    double* parts[6] = {dtmp, dtmp + maxl, dtmp + 2 * maxl, dtmp + 3 * maxl,
                        dtmp + 4 * maxl, dtmp + 5 * maxl};
    for (int k = 0; k < lx; ++k) hbuf[k] = width[i] / global_nx[i];
    HIPCHECK(hipMemcpy(parts[0], hbuf, sizeof(double) * lx, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(parts[2], hbuf, sizeof(double) * lx, hipMemcpyHostToDevice));
    for (int k = 0; k < ly; ++k) hbuf[k] = height[i] / global_ny[i];
    HIPCHECK(hipMemcpy(parts[1], hbuf, sizeof(double) * ly, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(parts[3], hbuf, sizeof(double) * ly, hipMemcpyHostToDevice));
    HIPCHECK(fk_reciprocal(lx, parts[4], parts[0], E->stream));
    HIPCHECK(fk_reciprocal(ly, parts[5], parts[1], E->stream));
    HIPCHECK(hipStreamSynchronize(E->stream));
    const int len[6] = {lx, ly, lx, ly, lx, ly};
    char* g = (char*)E->geo + gof[i];
    size_t gpos = 0;
    for (int k = 0; k < 6; ++k) {
      E->h_insts[i].geo[k] = g + gpos;
      if (fp32)
        HIPCHECK(fk_narrow(len[k], (float*)(g + gpos), parts[k], E->stream));
      else
        HIPCHECK(hipMemcpy(g + gpos, parts[k], sizeof(double) * len[k],
                           hipMemcpyDeviceToDevice));
      gpos += (size_t)len[k] * (size_t)E->esize;
    }
    HIPCHECK(hipStreamSynchronize(E->stream));
    FkEnsInst* in = &E->h_insts[i];
    in->nx = global_nx[i] + 4;
    in->ny = global_ny[i] + 4;
    for (int k = 0; k < 16; ++k) in->a[k] = (char*)E->pool + aoff[16 * i + k];
    in->dtp = E->d_dtp + 2 * i;
    in->max_dt = max_dt[i];
    E->h_dtp[2 * i] = C_T * dt_param[i];
    E->h_dtp[2 * i + 1] = C_T * dt_param[i];
  }
  HIPCHECK(hipMemcpy(E->d_dtp, E->h_dtp, sizeof(double) * 2 * (size_t)ninst,
                     hipMemcpyHostToDevice));
  /* uniform spacings: powers of two within the type's safe exponent range
   * let the kernels divide by them as products with exact reciprocals */
  E->pow2 = 1;
  for (int i = 0; i < ninst && E->pow2; ++i) {
    const double d[2] = {width[i] / global_nx[i], height[i] / global_ny[i]};
    for (int k = 0; k < 2; ++k) {
      int e = 0;
      const double f = frexp(d[k], &e);
      if (!(f == 0.5 && e - 1 <= (fp32 ? 40 : 300) && e - 1 >= -(fp32 ? 40 : 300)))
        E->pow2 = 0;
    }
  }
  {
    const char* p2 = getenv("FLOWHIP_POW2");
    if (p2 && !strcmp(p2, "0")) E->pow2 = 0;
  }
  /* every spacing array of an instance holds width / global_nx (or height /
   * global_ny) throughout; FLOWHIP_UNIFORM=0 keeps the per-row loads */
  {
    const char* un = getenv("FLOWHIP_UNIFORM");
    E->uniform = (un && !strcmp(un, "0")) ? 0 : 1;
  }
  HIPCHECK(hipFree(dtmp));
  free(hbuf);
  free(aoff);
  free(gof);
  HIPCHECK(hipMalloc((void**)&E->d_insts, sizeof(FkEnsInst) * (size_t)ninst));
  HIPCHECK(hipMalloc((void**)&E->hist.counter, sizeof(int)));
  HIPCHECK(hipMemset(E->hist.counter, 0, sizeof(int)));
  HIPCHECK(hipMalloc((void**)&E->d_queue, sizeof(unsigned int) * FK_MARCH_QUEUE_WORDS));
  HIPCHECK(hipMemset(E->d_queue, 0, sizeof(unsigned int) * FK_MARCH_QUEUE_WORDS));
  ens_build_work(E);
  return E;
}

void flowhip_ensemble_delete(FlowHipEnsemble* E) {
  if (!E) return;
  HIPCHECK(hipStreamSynchronize(E->stream));
  ens_drop_graphs(E);
  HIPCHECK(hipFree(E->pool));
  HIPCHECK(hipFree(E->snap));
  HIPCHECK(hipFree(E->geo));
  HIPCHECK(hipFree(E->d_dtp));
  HIPCHECK(hipHostFree(E->h_dtp));
  HIPCHECK(hipFree(E->d_work));
  HIPCHECK(hipFree(E->d_insts));
  HIPCHECK(hipFree(E->part.min));
  HIPCHECK(hipFree(E->part.nan));
  HIPCHECK(hipFree(E->part.sums));
  HIPCHECK(hipFree(E->hist.dt));
  HIPCHECK(hipFree(E->hist.sums));
  HIPCHECK(hipFree(E->hist.nan));
  HIPCHECK(hipFree(E->hist.counter));
  HIPCHECK(hipFree(E->d_queue));
  HIPCHECK(hipStreamDestroy(E->stream));
  free(E->gnx);
  free(E->gny);
  free(E->h_insts);
  free(E);
}

int flowhip_ensemble_size(FlowHipEnsemble* E) { return E ? E->ninst : 0; }

int flowhip_ensemble_local_size(FlowHipEnsemble* E, int inst, int* nx, int* ny) {
  if (!E || inst < 0 || inst >= E->ninst) return -1;
  if (nx) *nx = E->h_insts[inst].nx;
  if (ny) *ny = E->h_insts[inst].ny;
  return 0;
}

void* flowhip_ensemble_array(FlowHipEnsemble* E, int inst, int which) {
  if (!E || inst < 0 || inst >= E->ninst || which < 0 || which >= 16) return NULL;
  return E->h_insts[inst].a[which];
}

/* Density and energy of one instance from host FP64 values (rounded to
 * FP32 in an FP32 ensemble), then the reflective boundary of main.c:56-59. */
int flowhip_ensemble_set_state(FlowHipEnsemble* E, int inst, const double* rho,
                               const double* e) {
  if (!E || inst < 0 || inst >= E->ninst || !rho || !e) return -1;
  const FkEnsInst* in = &E->h_insts[inst];
  const size_t n = (size_t)(in->nx + 1) * (size_t)(in->ny + 1);
  const double* src[2] = {rho, e};
  void* dst[2] = {in->a[1], in->a[3]};
  for (int k = 0; k < 2; ++k) {
    if (E->fp32) {
      float* tmp = (float*)malloc(sizeof(float) * n);
      if (!tmp) TERMINATE("ensemble: out of host memory");
      for (size_t q = 0; q < n; ++q) tmp[q] = (float)src[k][q];
      HIPCHECK(hipMemcpy(dst[k], tmp, sizeof(float) * n, hipMemcpyHostToDevice));
      free(tmp);
      HIPCHECK(fk_boundary_f32(in->nx, in->ny, 2, (float*)dst[k], NO_INVERT, E->stream));
    } else {
      HIPCHECK(hipMemcpy(dst[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice));
      HIPCHECK(fk_boundary(in->nx, in->ny, 2, (double*)dst[k], NO_INVERT, E->stream));
    }
  }
  HIPCHECK(hipStreamSynchronize(E->stream));
  return 0;
}

/* One array of one instance, widened to FP64, into host memory of
 * (nx+1)*(ny+1) doubles. */
int flowhip_ensemble_get_array(FlowHipEnsemble* E, int inst, int which, double* host) {
  if (!E || inst < 0 || inst >= E->ninst || which < 0 || which >= 16 || !host)
    return -1;
  const FkEnsInst* in = &E->h_insts[inst];
  const size_t n = (size_t)(in->nx + 1) * (size_t)(in->ny + 1);
  HIPCHECK(hipStreamSynchronize(E->stream));
  if (E->fp32) {
    float* tmp = (float*)malloc(sizeof(float) * n);
    if (!tmp) TERMINATE("ensemble: out of host memory");
    HIPCHECK(hipMemcpy(tmp, in->a[which], sizeof(float) * n, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) host[q] = (double)tmp[q];
    free(tmp);
  } else {
    HIPCHECK(hipMemcpy(host, in->a[which], sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  return 0;
}

int flowhip_ensemble_set_dt(FlowHipEnsemble* E, int inst, double dt, double dt_h) {
  if (!E || inst < 0 || inst >= E->ninst) return -1;
  HIPCHECK(hipStreamSynchronize(E->stream));
  E->h_dtp[0] = dt;
  E->h_dtp[1] = dt_h;
  HIPCHECK(hipMemcpy(E->d_dtp + 2 * inst, E->h_dtp, 2 * sizeof(double),
                     hipMemcpyHostToDevice));
  return 0;
}

int flowhip_ensemble_set_graphs(FlowHipEnsemble* E, int on) {
  if (!E) return -1;
  E->use_graphs = on ? 1 : 0;
  return 0;
}

int flowhip_ensemble_set_schedule(FlowHipEnsemble* E, int chunk_rows, int order,
                                  int fused_momentum) {
  if (!E || chunk_rows < 4 || chunk_rows > 120 || order < 0 || order > 2) return -1;
  HIPCHECK(hipStreamSynchronize(E->stream));
  E->chunk_rows = chunk_rows;
  E->order = order;
  E->fused_momentum = fused_momentum ? 1 : 0;
  ens_build_work(E);
  return 0;
}

int flowhip_ensemble_get_schedule(FlowHipEnsemble* E, int* chunk_rows, int* order,
                                  int* fused_momentum) {
  if (!E) return -1;
  if (chunk_rows) *chunk_rows = E->chunk_rows;
  if (order) *order = E->order;
  if (fused_momentum) *fused_momentum = E->fused_momentum;
  return 0;
}

/* History buffers for rows steps; graphs hold their addresses. */
static void ens_reserve_history(FlowHipEnsemble* E, int rows) {
  if (rows <= E->hist_rows) return;
  HIPCHECK(hipStreamSynchronize(E->stream));
  HIPCHECK(hipFree(E->hist.dt));
  HIPCHECK(hipFree(E->hist.sums));
  HIPCHECK(hipFree(E->hist.nan));
  const size_t n = (size_t)rows * (size_t)E->ninst;
  HIPCHECK(hipMalloc((void**)&E->hist.dt, sizeof(double) * 2 * n));
  HIPCHECK(hipMalloc((void**)&E->hist.sums, sizeof(double) * 2 * n));
  HIPCHECK(hipMalloc((void**)&E->hist.nan, sizeof(int) * n));
  E->hist_rows = rows;
  ens_drop_graphs(E);
}

/* One step; replayed from the graph of its sweep parity unless direct
 * (the first step of a run, which also has tt == 0's first_step dt). */
static void ens_step(FlowHipEnsemble* E, int tt, int direct) {
  const int parity = tt & 1;
  if (E->use_graphs && !direct && tt != 0) {
    if (!E->graph_ok[parity]) {
      hipGraph_t g;
      HIPCHECK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
      HIPCHECK(fk_ens_step(ens_launch(E), tt, E->fp32, E->stream));
      HIPCHECK(hipStreamEndCapture(E->stream, &g));
      HIPCHECK(hipGraphInstantiate(&E->graph[parity], g, NULL, NULL, 0));
      HIPCHECK(hipGraphDestroy(g));
      E->graph_ok[parity] = 1;
    }
    HIPCHECK(hipGraphLaunch(E->graph[parity], E->stream));
  } else {
    HIPCHECK(fk_ens_step(ens_launch(E), tt, E->fp32, E->stream));
  }
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* main.c:64-115 for every instance at once, without sim_end (every
 * instance takes niters steps). dt_pairs: [instance][niters+1][2] (the
 * pre-loop pair, then after each step); mass, energy, undefined:
 * [instance][niters]; any may be NULL. Returns niters. */
int flowhip_ensemble_run(FlowHipEnsemble* E, int niters, int tt0, int pre_timestep,
                         double* dt_pairs, double* mass, double* energy,
                         double* wall_seconds, int* undefined) {
  if (!E || niters < 0) return -1;
  const int n = E->ninst;
  ens_reserve_history(E, niters > 0 ? niters : 1);
  if (pre_timestep) HIPCHECK(fk_ens_timestep(ens_launch(E), E->fp32, E->stream));
  HIPCHECK(hipMemcpyAsync(E->h_dtp, E->d_dtp, sizeof(double) * 2 * (size_t)n,
                          hipMemcpyDeviceToHost, E->stream));
  HIPCHECK(fk_ens_set_counter(E->hist.counter, 0, E->stream));
  HIPCHECK(hipStreamSynchronize(E->stream));
  if (dt_pairs)
    for (int i = 0; i < n; ++i) {
      dt_pairs[(size_t)i * 2 * (niters + 1)] = E->h_dtp[2 * i];
      dt_pairs[(size_t)i * 2 * (niters + 1) + 1] = E->h_dtp[2 * i + 1];
    }
  const double t0 = now_seconds();
  for (int k = 0; k < niters; ++k) ens_step(E, tt0 + k, k == 0);
  HIPCHECK(hipStreamSynchronize(E->stream));
  if (wall_seconds) *wall_seconds = now_seconds() - t0;
  if (niters > 0 && (dt_pairs || mass || energy)) {
    const size_t m = (size_t)niters * (size_t)n;
    double* hd = (double*)malloc(sizeof(double) * 2 * m);
    double* hs = (double*)malloc(sizeof(double) * 2 * m);
    if (!hd || !hs) TERMINATE("ensemble: out of host memory");
    HIPCHECK(hipMemcpy(hd, E->hist.dt, sizeof(double) * 2 * m, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hs, E->hist.sums, sizeof(double) * 2 * m, hipMemcpyDeviceToHost));
    for (int k = 0; k < niters; ++k)
      for (int i = 0; i < n; ++i) {
        const size_t h = (size_t)k * n + i;
        if (dt_pairs) {
          dt_pairs[(size_t)i * 2 * (niters + 1) + 2 * (k + 1)] = hd[2 * h];
          dt_pairs[(size_t)i * 2 * (niters + 1) + 2 * (k + 1) + 1] = hd[2 * h + 1];
        }
        if (mass) mass[(size_t)i * niters + k] = hs[2 * h];
        if (energy) energy[(size_t)i * niters + k] = hs[2 * h + 1];
      }
    /* a diverged instance is fatal, as in the single-mesh form */
    for (int i = 0; i < n; ++i) {
      const size_t h = (size_t)(niters - 1) * n + i;
      if (!isfinite(hd[2 * h]) || !(hd[2 * h] > 0.0))
        TERMINATE("ensemble: instance %d has a non-finite or non-positive "
                  "timestep (dt=%.12e): its state has diverged", i, hd[2 * h]);
    }
    free(hd);
    free(hs);
  }
  if (niters > 0 && undefined) {
    const size_t m = (size_t)niters * (size_t)n;
    int* hn = (int*)malloc(sizeof(int) * m);
    if (!hn) TERMINATE("ensemble: out of host memory");
    HIPCHECK(hipMemcpy(hn, E->hist.nan, sizeof(int) * m, hipMemcpyDeviceToHost));
    for (int k = 0; k < niters; ++k)
      for (int i = 0; i < n; ++i) undefined[(size_t)i * niters + k] = hn[(size_t)k * n + i];
    free(hn);
  }
  return niters;
}

/* Try every chunk height in {16, 20, 24, 32, 48, 64}, work order (instance,
 * largest first, round robin) and momentum schedule (fused, x then y) for
 * trial_steps steps from the current state; keep the fastest and put the
 * state back. */
int flowhip_ensemble_autotune(FlowHipEnsemble* E, int trial_steps, int tt0,
                              double* best_ms) {
  if (!E || trial_steps < 1) return -1;
  const int n = E->ninst;
  if (!E->snap) HIPCHECK(hipMalloc(&E->snap, E->pool_bytes + sizeof(double) * 2 * (size_t)n));
  HIPCHECK(hipStreamSynchronize(E->stream));
  HIPCHECK(hipMemcpy(E->snap, E->pool, E->pool_bytes, hipMemcpyDeviceToDevice));
  HIPCHECK(hipMemcpy((char*)E->snap + E->pool_bytes, E->d_dtp,
                     sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToDevice));
  const int rows[] = {16, 20, 24, 32, 48, 64};
  double best = 1e300;
  int brow = E->chunk_rows, bord = E->order, bfm = E->fused_momentum;
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  ens_reserve_history(E, trial_steps + 1);
  for (int r = 0; r < (int)(sizeof(rows) / sizeof(rows[0])); ++r)
    for (int o = 0; o < 3; ++o)
    for (int fm = 0; fm < 2; ++fm) {
      E->chunk_rows = rows[r];
      E->order = o;
      E->fused_momentum = fm;
      ens_build_work(E);
      HIPCHECK(fk_ens_set_counter(E->hist.counter, 0, E->stream));
      /* one untimed step, then the timed ones */
      HIPCHECK(fk_ens_step(ens_launch(E), tt0 + 1, E->fp32, E->stream));
      HIPCHECK(hipEventRecord(a, E->stream));
      for (int k = 0; k < trial_steps; ++k)
        HIPCHECK(fk_ens_step(ens_launch(E), tt0 + 2 + k, E->fp32, E->stream));
      HIPCHECK(hipEventRecord(b, E->stream));
      HIPCHECK(hipEventSynchronize(b));
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, a, b));
      if (ms < best) {
        best = ms;
        brow = rows[r];
        bord = o;
        bfm = fm;
      }
      HIPCHECK(hipMemcpyAsync(E->pool, E->snap, E->pool_bytes, hipMemcpyDeviceToDevice,
                              E->stream));
      HIPCHECK(hipMemcpyAsync(E->d_dtp, (char*)E->snap + E->pool_bytes,
                              sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToDevice,
                              E->stream));
      HIPCHECK(hipStreamSynchronize(E->stream));
    }
  HIPCHECK(hipEventDestroy(a));
  HIPCHECK(hipEventDestroy(b));
  E->chunk_rows = brow;
  E->order = bord;
  E->fused_momentum = bfm;
  ens_build_work(E);
  if (best_ms) *best_ms = best / trial_steps;
  return 0;
}
