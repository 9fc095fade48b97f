/*
 * ctx.c -- creation of the hip kernel set's per-Mesh state and its
 * per-kernel event timers.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "flowhip_ctx.h"

/* FLOWHIP_SCHEDULE (stagewise | fused | march) and FLOWHIP_MARCH_ROWS pick
 * the schedule for callers that only see flow_interface.h, such as flow's
 * own driver. */
static void schedule_from_env(FlowHipCtx* h) {
  const char* s = getenv("FLOWHIP_SCHEDULE");
  if (s && *s) {
    if (!strcmp(s, "stagewise"))
      h->schedule = FLOWHIP_STAGEWISE;
    else if (!strcmp(s, "fused"))
      h->schedule = FLOWHIP_FUSED;
    else if (!strcmp(s, "march"))
      h->schedule = FLOWHIP_MARCH;
    else
      TERMINATE("FLOWHIP_SCHEDULE must be stagewise, fused or march, not '%s'", s);
  }
  const char* c = getenv("FLOWHIP_CFL_RULE");
  if (c && *c) {
    if (!strcmp(c, "defined"))
      h->cfl_rule = FLOWHIP_CFL_DEFINED;
    else if (!strcmp(c, "omp3-serial"))
      h->cfl_rule = FLOWHIP_CFL_OMP3_SERIAL;
    else
      TERMINATE("FLOWHIP_CFL_RULE must be defined or omp3-serial, not '%s'", c);
  }
  const char* r = getenv("FLOWHIP_MARCH_ROWS");
  if (r && *r) {
    char* end = NULL;
    const long rows = strtol(r, &end, 10);
    if (*end != '\0' || rows < 4 || rows > 120)
      TERMINATE("FLOWHIP_MARCH_ROWS must be an integer in [4, 120], not '%s'", r);
    h->march_rows = (int)rows;
  }
}

/* Live contexts, for flowhip_print_profile. */
#define FLOWHIP_MAX_LIVE 64
static FlowHipCtx* g_live[FLOWHIP_MAX_LIVE];

static void live_add(FlowHipCtx* h) {
  for (int k = 0; k < FLOWHIP_MAX_LIVE; ++k)
    if (!g_live[k]) {
      g_live[k] = h;
      return;
    }
}

static void live_remove(FlowHipCtx* h) {
  for (int k = 0; k < FLOWHIP_MAX_LIVE; ++k)
    if (g_live[k] == h) g_live[k] = NULL;
}

FlowHipCtx* flowhip_ctx(Mesh* mesh) {
  if (mesh->hip) return (FlowHipCtx*)mesh->hip;
  FlowHipCtx* h = (FlowHipCtx*)calloc(1, sizeof(FlowHipCtx));
  if (!h) TERMINATE("hip kernel set: out of host memory");
  HIPCHECK(hipGetDevice(&h->device));
  /* default priority: the kernel set never uses prioritised streams */
  HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const int np = fk_reduce_blocks();
  HIPCHECK(hipMalloc((void**)&h->red.partials, sizeof(double) * np));
  HIPCHECK(hipMalloc((void**)&h->red.ticket, sizeof(unsigned int) * 4));
  HIPCHECK(hipMalloc((void**)&h->red.dtp, sizeof(double) * 2));
  HIPCHECK(hipMalloc((void**)&h->red.sums, sizeof(double) * 4));
  HIPCHECK(hipMalloc((void**)&h->red.status, sizeof(int) * 4));
  HIPCHECK(hipMalloc((void**)&h->red.queue, sizeof(unsigned int) * FK_MARCH_QUEUE_WORDS));
  HIPCHECK(hipMalloc((void**)&h->red.totals,
                     sizeof(unsigned long long) * FK_TOTALS_WORDS));
  HIPCHECK(hipMemset(h->red.partials, 0, sizeof(double) * np));
  HIPCHECK(hipMemset(h->red.ticket, 0, sizeof(unsigned int) * 4));
  HIPCHECK(hipMemset(h->red.dtp, 0, sizeof(double) * 2));
  HIPCHECK(hipMemset(h->red.sums, 0, sizeof(double) * 4));
  HIPCHECK(hipMemset(h->red.status, 0, sizeof(int) * 4));
  HIPCHECK(hipMemset(h->red.queue, 0, sizeof(unsigned int) * FK_MARCH_QUEUE_WORDS));
  HIPCHECK(hipMemset(h->red.totals, 0, sizeof(unsigned long long) * FK_TOTALS_WORDS));
  h->red.max_blocks = np / 4;
  HIPCHECK(hipHostMalloc((void**)&h->h_stage, sizeof(double) * 8,
                         hipHostMallocDefault));
  HIPCHECK(hipHostMalloc((void**)&h->h_status, sizeof(int) * 4,
                         hipHostMallocDefault));
  memset(h->h_stage, 0, sizeof(double) * 8);
  memset(h->h_status, 0, sizeof(int) * 4);
  h->schedule = FLOWHIP_MARCH;
  h->march_rows = 0; /* each march kernel's own chunk height */
  schedule_from_env(h);
  /* FLOWHIP_POW2=0 keeps IEEE quotients by spacings even on power-of-two
   * meshes (the results are the same bits either way) */
  const char* p2 = getenv("FLOWHIP_POW2");
  h->no_pow2 = (p2 && !strcmp(p2, "0")) ? 1 : 0;
  /* FLOWHIP_UNIFORM=0 keeps the per-row geometry loads on uniform meshes
   * (the same bits either way) */
  const char* un = getenv("FLOWHIP_UNIFORM");
  h->no_uniform = (un && !strcmp(un, "0")) ? 1 : 0;
  /* two columns per lane in the Lagrangian and mass/energy marches (the
   * same fields either way): FLOWHIP_PAIRS=0 never, =1 on every mesh,
   * unset on meshes of at least FK_PAIR_MIN_CELLS interior cells */
  const char* pr = getenv("FLOWHIP_PAIRS");
  h->pairs = !pr ? 1 : !strcmp(pr, "0") ? 0 : !strcmp(pr, "1") ? 2 : 1;
  HIPCHECK(hipDeviceSynchronize());
  mesh->hip = h;
  live_add(h);
  const char* prof = getenv("FLOWHIP_PROFILE");
  if (prof && !strcmp(prof, "1")) flowhip_set_timing(mesh, 1);
  return h;
}

void flowhip_ctx_destroy(Mesh* mesh) {
  FlowHipCtx* h = (FlowHipCtx*)mesh->hip;
  if (!h) return;
  live_remove(h);
  hipStreamSynchronize(h->stream);
  for (int g = 0; g < 3; ++g)
    if (h->graph_ok[g]) hipGraphExecDestroy(h->graph[g]);
  if (h->events_created) {
    for (int k = 0; k < FLOWHIP_MAX_PENDING; ++k) {
      hipEventDestroy(h->ev_start[k]);
      hipEventDestroy(h->ev_stop[k]);
    }
  }
  hipFree(h->inv_edgedx);
  hipFree(h->inv_edgedy);
  hipFree(h->geo32);
  hipFree(h->cfl_buf);
  hipFree(h->red.partials);
  hipFree(h->red.ticket);
  hipFree(h->red.dtp);
  hipFree(h->red.sums);
  hipFree(h->red.status);
  hipFree(h->red.queue);
  hipFree(h->red.totals);
  hipHostFree(h->h_stage);
  hipHostFree(h->h_status);
  hipStreamDestroy(h->stream);
  free(h);
  mesh->hip = NULL;
}

/* Whether every value is a power of two with |exponent| <= emax. */
static int all_pow2(const double* dev, int n, int emax) {
  double* host = (double*)malloc(sizeof(double) * (size_t)n);
  if (!host) TERMINATE("hip kernel set: out of host memory");
  HIPCHECK(hipMemcpy(host, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  int ok = 1;
  for (int k = 0; k < n && ok; ++k) {
    int e = 0;
    const double f = frexp(host[k], &e);
    ok = (f == 0.5) && (e - 1 <= emax) && (e - 1 >= -emax);
  }
  free(host);
  return ok;
}

/* Whether all n values are the same. */
static int all_equal(const double* dev, int n) {
  double* host = (double*)malloc(sizeof(double) * (size_t)n);
  if (!host) TERMINATE("hip kernel set: out of host memory");
  HIPCHECK(hipMemcpy(host, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  int ok = 1;
  for (int k = 1; k < n && ok; ++k) ok = (host[k] == host[0]);
  free(host);
  return ok;
}

/* 1/edgedx and 1/edgedy, recomputed when the geometry arrays change, and
 * whether the spacings are powers of two and uniform. */
static void ensure_reciprocals(FlowHipCtx* h, const Mesh* mesh) {
  const int lx = mesh->local_nx + 1, ly = mesh->local_ny + 1;
  if (h->inv_src_x == mesh->edgedx && h->inv_src_y == mesh->edgedy &&
      h->inv_src_cx == mesh->celldx && h->inv_src_cy == mesh->celldy &&
      h->inv_nx == lx && h->inv_ny == ly)
    return;
  HIPCHECK(hipStreamSynchronize(h->stream));
  /* FP64: products of up to three spacings or reciprocals stay normal; the
   * FP32 form's narrower range asks for smaller exponents */
  h->pow2 = all_pow2(mesh->edgedx, lx, 300) && all_pow2(mesh->edgedy, ly, 300) &&
            all_pow2(mesh->celldx, lx, 300) && all_pow2(mesh->celldy, ly, 300);
  h->pow2_f32 = h->pow2 && all_pow2(mesh->edgedx, lx, 40) &&
                all_pow2(mesh->edgedy, ly, 40) && all_pow2(mesh->celldx, lx, 40) &&
                all_pow2(mesh->celldy, ly, 40);
  h->uniform = all_equal(mesh->edgedx, lx) && all_equal(mesh->edgedy, ly) &&
               all_equal(mesh->celldx, lx) && all_equal(mesh->celldy, ly);
  h->inv_src_cx = mesh->celldx;
  h->inv_src_cy = mesh->celldy;
  hipFree(h->inv_edgedx);
  hipFree(h->inv_edgedy);
  HIPCHECK(hipMalloc((void**)&h->inv_edgedx, sizeof(double) * lx));
  HIPCHECK(hipMalloc((void**)&h->inv_edgedy, sizeof(double) * ly));
  HIPCHECK(fk_reciprocal(lx, h->inv_edgedx, mesh->edgedx, h->stream));
  HIPCHECK(fk_reciprocal(ly, h->inv_edgedy, mesh->edgedy, h->stream));
  h->inv_src_x = mesh->edgedx;
  h->inv_src_y = mesh->edgedy;
  h->inv_nx = lx;
  h->inv_ny = ly;
}

FkMesh flowhip_fkmesh(const Mesh* mesh) {
  FkMesh m;
  FlowHipCtx* h = (FlowHipCtx*)mesh->hip;
  if (h) {
    ensure_reciprocals(h, mesh);
    m.inv_edgedx = h->inv_edgedx;
    m.inv_edgedy = h->inv_edgedy;
    m.pow2 = h->pow2 && !h->no_pow2;
    m.uniform = m.pow2 && h->uniform && !h->no_uniform;
    m.pairs = h->pairs;
  } else {
    m.inv_edgedx = m.inv_edgedy = NULL;
    m.pow2 = 0;
    m.uniform = 0;
    m.pairs = 0;
  }
  m.nx = mesh->local_nx;
  m.ny = mesh->local_ny;
  m.pad = mesh->pad;
  m.edgedx = mesh->edgedx;
  m.edgedy = mesh->edgedy;
  m.celldx = mesh->celldx;
  m.celldy = mesh->celldy;
  return m;
}

void flowhip_invalidate_sums(Mesh* mesh) {
  if (mesh && mesh->hip) ((FlowHipCtx*)mesh->hip)->sums_valid = 0;
}

FkReduce flowhip_lagrangian_red(FlowHipCtx* h, int nx, int ny) {
  FkReduce r = h->red;
  r.cfl = NULL;
  if (h->cfl_rule != FLOWHIP_CFL_OMP3_SERIAL) return r;
  const size_t need = (size_t)nx * (size_t)ny;
  if (h->cfl_len < need) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    hipFree(h->cfl_buf);
    h->cfl_buf = NULL;
    h->cfl_len = 0;
    HIPCHECK(hipMalloc((void**)&h->cfl_buf, sizeof(double) * need));
    h->cfl_len = need;
  }
  r.cfl = h->cfl_buf;
  return r;
}

int flowhip_set_cfl_rule(Mesh* mesh, int rule) {
  if (rule != FLOWHIP_CFL_DEFINED && rule != FLOWHIP_CFL_OMP3_SERIAL) return -1;
  flowhip_ctx(mesh)->cfl_rule = rule;
  return 0;
}

int flowhip_get_cfl_rule(Mesh* mesh) { return flowhip_ctx(mesh)->cfl_rule; }

int flowhip_undefined_candidates(Mesh* mesh) { return flowhip_ctx(mesh)->h_status[1]; }

static int timer_index(FlowHipCtx* h, const char* name) {
  for (int k = 0; k < h->ntimers; ++k)
    if (h->timers[k].name == name || strcmp(h->timers[k].name, name) == 0)
      return k;
  if (h->ntimers == FLOWHIP_MAX_TIMERS) return -1;
  h->timers[h->ntimers].name = name;
  h->timers[h->ntimers].total_ms = 0.0;
  h->timers[h->ntimers].launches = 0;
  return h->ntimers++;
}

int flowhip_timer_begin(FlowHipCtx* h, const char* name) {
  if (!h->timing) return -1;
  if (h->npending == FLOWHIP_MAX_PENDING) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    flowhip_timer_collect(h);
  }
  const int t = timer_index(h, name);
  if (t < 0) return -1;
  const int slot = h->npending++;
  h->pending_timer[slot] = t;
  HIPCHECK(hipEventRecord(h->ev_start[slot], h->stream));
  return slot;
}

void flowhip_timer_end(FlowHipCtx* h, int slot) {
  if (slot < 0) return;
  HIPCHECK(hipEventRecord(h->ev_stop[slot], h->stream));
}

void flowhip_timer_collect(FlowHipCtx* h) {
  for (int k = 0; k < h->npending; ++k) {
    float ms = 0.0f;
    HIPCHECK(hipEventSynchronize(h->ev_stop[k]));
    HIPCHECK(hipEventElapsedTime(&ms, h->ev_start[k], h->ev_stop[k]));
    FlowHipTimer* t = &h->timers[h->pending_timer[k]];
    t->total_ms += ms;
    t->launches += 1;
  }
  h->npending = 0;
}

int flowhip_set_timing(Mesh* mesh, int on) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  if (on && !h->events_created) {
    for (int k = 0; k < FLOWHIP_MAX_PENDING; ++k) {
      HIPCHECK(hipEventCreate(&h->ev_start[k]));
      HIPCHECK(hipEventCreate(&h->ev_stop[k]));
    }
    h->events_created = 1;
  }
  if (!on && h->npending) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    flowhip_timer_collect(h);
  }
  h->timing = on ? 1 : 0;
  return 0;
}

int flowhip_timing_count(Mesh* mesh) { return flowhip_ctx(mesh)->ntimers; }

int flowhip_timing_entry(Mesh* mesh, int k, char* name, int name_len,
                         double* total_ms, long* launches) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  if (h->npending) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    flowhip_timer_collect(h);
  }
  if (k < 0 || k >= h->ntimers) return -1;
  if (name && name_len > 0) {
    strncpy(name, h->timers[k].name, (size_t)name_len - 1);
    name[name_len - 1] = '\0';
  }
  if (total_ms) *total_ms = h->timers[k].total_ms;
  if (launches) *launches = h->timers[k].launches;
  return 0;
}

int flowhip_timing_reset(Mesh* mesh) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  if (h->npending) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    flowhip_timer_collect(h);
  }
  h->ntimers = 0;
  return 0;
}

/* The geometry rounded to FP32, rebuilt when the mesh's arrays change. */
FkMesh32 flowhip_fkmesh32(Mesh* mesh) {
  FlowHipCtx* h = flowhip_ctx(mesh);
  const FkMesh m = flowhip_fkmesh(mesh);
  const int lx = mesh->local_nx + 1, ly = mesh->local_ny + 1;
  if (!(h->geo32 && h->geo32_src[0] == mesh->edgedx &&
        h->geo32_src[1] == mesh->edgedy && h->geo32_src[2] == mesh->celldx &&
        h->geo32_src[3] == mesh->celldy && h->geo32_nx == lx &&
        h->geo32_ny == ly)) {
    hipFree(h->geo32);
    HIPCHECK(hipMalloc((void**)&h->geo32, sizeof(float) * 3 * (lx + ly)));
    float* g = h->geo32;
    HIPCHECK(fk_narrow(lx, g, m.edgedx, h->stream));
    HIPCHECK(fk_narrow(ly, g + lx, m.edgedy, h->stream));
    HIPCHECK(fk_narrow(lx, g + lx + ly, m.celldx, h->stream));
    HIPCHECK(fk_narrow(ly, g + 2 * lx + ly, m.celldy, h->stream));
    HIPCHECK(fk_narrow(lx, g + 2 * (lx + ly), m.inv_edgedx, h->stream));
    HIPCHECK(fk_narrow(ly, g + 3 * lx + 2 * ly, m.inv_edgedy, h->stream));
    h->geo32_src[0] = mesh->edgedx;
    h->geo32_src[1] = mesh->edgedy;
    h->geo32_src[2] = mesh->celldx;
    h->geo32_src[3] = mesh->celldy;
    h->geo32_nx = lx;
    h->geo32_ny = ly;
  }
  FkMesh32 f;
  f.nx = m.nx;
  f.ny = m.ny;
  f.pad = m.pad;
  f.edgedx = h->geo32;
  f.edgedy = h->geo32 + lx;
  f.celldx = h->geo32 + lx + ly;
  f.celldy = h->geo32 + 2 * lx + ly;
  f.inv_edgedx = h->geo32 + 2 * (lx + ly);
  f.inv_edgedy = h->geo32 + 3 * lx + 2 * ly;
  f.pow2 = h->pow2_f32 && !h->no_pow2;
  f.uniform = f.pow2 && h->uniform && !h->no_uniform;
  f.pairs = h->pairs;
  return f;
}

/* arch's PRINT_PROFILING_RESULTS: the hipEvent times of every live mesh's
 * kernels (meshes created with FLOWHIP_PROFILE=1), longest first. */
void flowhip_print_profile(const struct Profile* p) {
  if (p == &comms_profile) {
    printf("Communication: none (one GPU; the reflective boundary and the "
           "reductions run on the device)\n");
    return;
  }
  int any = 0;
  for (int m = 0; m < FLOWHIP_MAX_LIVE; ++m) {
    FlowHipCtx* h = g_live[m];
    if (!h || !h->ntimers) continue;
    if (h->npending) {
      HIPCHECK(hipStreamSynchronize(h->stream));
      flowhip_timer_collect(h);
    }
    int order[FLOWHIP_MAX_TIMERS];
    for (int k = 0; k < h->ntimers; ++k) order[k] = k;
    for (int a = 1; a < h->ntimers; ++a)
      for (int b = a; b > 0 && h->timers[order[b]].total_ms >
                                   h->timers[order[b - 1]].total_ms; --b) {
        const int t = order[b];
        order[b] = order[b - 1];
        order[b - 1] = t;
      }
    double total = 0.0;
    for (int k = 0; k < h->ntimers; ++k) total += h->timers[k].total_ms;
    printf("\nKernel times (hipEvents on the mesh's stream):\n");
    printf("%-24s %10s %14s %14s %8s\n", "kernel", "launches", "total (ms)",
           "per launch (ms)", "share");
    for (int k = 0; k < h->ntimers; ++k) {
      const FlowHipTimer* t = &h->timers[order[k]];
      printf("%-24s %10ld %14.3f %14.4f %7.1f%%\n", t->name, t->launches, t->total_ms,
             t->launches ? t->total_ms / (double)t->launches : 0.0,
             total > 0.0 ? 100.0 * t->total_ms / total : 0.0);
    }
    any = 1;
  }
  if (!any)
    printf("Kernel times: none recorded (set FLOWHIP_PROFILE=1 to time the "
           "kernels)\n");
}
