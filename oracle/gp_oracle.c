/*
 * gp_oracle.c -- CPU oracle for the DEM compile path.  TEST INFRASTRUCTURE
 * ONLY: see gp_oracle.h for what it restates and who may call it.
 *
 * Plain C11, single-threaded, FP64 throughout.
 */
#define _GNU_SOURCE
#include "gp_oracle.h"

#include <ctype.h>
#include <math.h>
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NONE32 0xFFFFFFFFu

/* ------------------------------------------------------------------ */
/* errors                                                               */
/* ------------------------------------------------------------------ */

static __thread char g_err[512];
static __thread jmp_buf *g_jmp;
static __thread int g_code;

const char *gpo_last_error(void) { return g_err; }

static void fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  g_code = code;
  longjmp(*g_jmp, 1);
}

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) fail(2, "out of memory");
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) fail(2, "out of memory");
  return p;
}
static void *xrealloc(void *p, size_t n) {
  void *q = realloc(p, n ? n : 1);
  if (!q) fail(2, "out of memory");
  return q;
}

/* ------------------------------------------------------------------ */
/* stim-text parser (gate subset of Circuit.hpp:425-462, REPEAT flattened  */
/* the way stim's for_each_operation does, Circuit.hpp:476-477)           */
/* ------------------------------------------------------------------ */

enum { K_TICK, K_R, K_M, K_CX, K_H, K_DEP1, K_DEP2, K_XERR, K_DET, K_OBS };

typedef struct {
  int kind;
  double arg;      /* probability, or observable index for K_OBS */
  uint32_t t0, nt; /* range in the target pool; rec targets stored as -k */
} op_t;

typedef struct {
  op_t *ops;
  size_t nops, capops;
  int32_t *tg;
  size_t ntg, captg;
} oplist_t;

typedef struct {
  oplist_t L;
  uint32_t nq, nticks, nmeas, ndet, nobs;
} circ_t;

static void ol_push_op(oplist_t *L, op_t o) {
  if (L->nops == L->capops) {
    L->capops = L->capops ? 2 * L->capops : 256;
    L->ops = xrealloc(L->ops, L->capops * sizeof(op_t));
  }
  L->ops[L->nops++] = o;
}
static void ol_push_tg(oplist_t *L, int32_t t) {
  if (L->ntg == L->captg) {
    L->captg = L->captg ? 2 * L->captg : 1024;
    L->tg = xrealloc(L->tg, L->captg * sizeof(int32_t));
  }
  L->tg[L->ntg++] = t;
}
static void ol_free(oplist_t *L) {
  free(L->ops);
  free(L->tg);
  memset(L, 0, sizeof *L);
}
static void ol_append(oplist_t *dst, const oplist_t *src) {
  for (size_t i = 0; i < src->nops; ++i) {
    op_t o = src->ops[i];
    uint32_t t0 = (uint32_t)dst->ntg;
    for (uint32_t k = 0; k < o.nt; ++k) ol_push_tg(dst, src->tg[o.t0 + k]);
    o.t0 = t0;
    ol_push_op(dst, o);
  }
}

typedef struct {
  const char *s;
  size_t pos, len;
  uint32_t line;
} lex_t;

/* Reads the next logical line into buf (comments stripped, trimmed). */
static int next_line(lex_t *lx, char **buf, size_t *cap) {
  while (lx->pos < lx->len) {
    size_t st = lx->pos;
    while (lx->pos < lx->len && lx->s[lx->pos] != '\n') lx->pos++;
    size_t en = lx->pos;
    if (lx->pos < lx->len) lx->pos++;
    lx->line++;
    for (size_t i = st; i < en; ++i)
      if (lx->s[i] == '#') {
        en = i;
        break;
      }
    while (st < en && isspace((unsigned char)lx->s[st])) st++;
    while (en > st && isspace((unsigned char)lx->s[en - 1])) en--;
    if (en == st) continue;
    if (en - st + 1 > *cap) {
      *cap = en - st + 1;
      *buf = xrealloc(*buf, *cap);
    }
    memcpy(*buf, lx->s + st, en - st);
    (*buf)[en - st] = 0;
    return 1;
  }
  return 0;
}

static int name_is(const char *a, const char *b) { return strcasecmp(a, b) == 0; }

static void parse_block(lex_t *lx, oplist_t *out, int depth, char **buf,
                        size_t *cap);

static uint32_t parse_uint(const char *s, uint32_t line) {
  char *e;
  if (!isdigit((unsigned char)*s)) fail(1, "line %u: bad target '%s'", line, s);
  unsigned long long v = strtoull(s, &e, 10);
  if (*e || v > 0x7FFFFFFFull) fail(1, "line %u: bad target '%s'", line, s);
  return (uint32_t)v;
}

static void parse_instruction(lex_t *lx, oplist_t *out, char *ln, int depth,
                              char **buf, size_t *cap) {
  uint32_t line = lx->line;
  char name[64];
  size_t i = 0;
  while (ln[i] && (isalnum((unsigned char)ln[i]) || ln[i] == '_') && i < 63) {
    name[i] = ln[i];
    i++;
  }
  name[i] = 0;
  if (i == 0) fail(1, "line %u: cannot parse '%s'", line, ln);
  char *p = ln + i;
  double args[16];
  int nargs = 0;
  while (*p == ' ' || *p == '\t') p++;
  if (*p == '(') {
    p++;
    for (;;) {
      while (isspace((unsigned char)*p)) p++;
      if (*p == ')') {
        p++;
        break;
      }
      char *e;
      double v = strtod(p, &e);
      if (e == p) fail(1, "line %u: bad argument in '%s'", line, ln);
      if (nargs < 16) args[nargs++] = v;
      p = e;
      while (isspace((unsigned char)*p)) p++;
      if (*p == ',') p++;
    }
  }
  if (name_is(name, "REPEAT")) {
    while (isspace((unsigned char)*p)) p++;
    char *e;
    unsigned long long reps = strtoull(p, &e, 10);
    if (e == p) fail(1, "line %u: bad REPEAT count", line);
    p = e;
    while (isspace((unsigned char)*p)) p++;
    if (*p != '{') fail(1, "line %u: REPEAT needs '{'", line);
    oplist_t body = {0};
    parse_block(lx, &body, depth + 1, buf, cap);
    for (unsigned long long r = 0; r < reps; ++r) ol_append(out, &body);
    ol_free(&body);
    return;
  }

  int kinds[2];
  int nk = 1;
  if (name_is(name, "TICK")) kinds[0] = K_TICK;
  else if (name_is(name, "R") || name_is(name, "RZ")) kinds[0] = K_R;
  else if (name_is(name, "M") || name_is(name, "MZ")) kinds[0] = K_M;
  else if (name_is(name, "MR") || name_is(name, "MRZ")) {
    kinds[0] = K_M;
    kinds[1] = K_R;
    nk = 2;
  } else if (name_is(name, "CX") || name_is(name, "CNOT") || name_is(name, "ZCX"))
    kinds[0] = K_CX;
  else if (name_is(name, "H") || name_is(name, "H_XZ")) kinds[0] = K_H;
  else if (name_is(name, "DEPOLARIZE1")) kinds[0] = K_DEP1;
  else if (name_is(name, "DEPOLARIZE2")) kinds[0] = K_DEP2;
  else if (name_is(name, "X_ERROR")) kinds[0] = K_XERR;
  else if (name_is(name, "DETECTOR")) kinds[0] = K_DET;
  else if (name_is(name, "OBSERVABLE_INCLUDE")) kinds[0] = K_OBS;
  else if (name_is(name, "QUBIT_COORDS") || name_is(name, "SHIFT_COORDS"))
    return; /* annotations: no effect on the error model */
  else
    fail(1, "Unsupported operation: %s", name);

  /* targets */
  uint32_t t0 = (uint32_t)out->ntg;
  char *tok = p;
  while (*tok) {
    while (isspace((unsigned char)*tok)) tok++;
    if (!*tok) break;
    char *e = tok;
    while (*e && !isspace((unsigned char)*e)) e++;
    char save = *e;
    *e = 0;
    if (kinds[0] == K_DET || kinds[0] == K_OBS) {
      if (strncasecmp(tok, "rec[-", 5) != 0) fail(1, "line %u: bad record target '%s'", line, tok);
      char *q = tok + 5;
      char *qe;
      unsigned long long k = strtoull(q, &qe, 10);
      if (qe == q || *qe != ']' || qe[1] != 0 || k == 0 || k > 0x7FFFFFFFull)
        fail(1, "line %u: bad record target '%s'", line, tok);
      ol_push_tg(out, -(int32_t)k);
    } else {
      ol_push_tg(out, (int32_t)parse_uint(tok, line));
    }
    *e = save;
    tok = e;
  }
  uint32_t nt = (uint32_t)out->ntg - t0;
  for (int k = 0; k < nk; ++k) {
    op_t o;
    o.kind = kinds[k];
    o.arg = 0.0;
    if (o.kind == K_M || o.kind == K_DEP1 || o.kind == K_DEP2 || o.kind == K_XERR) {
      if (o.kind != K_M && nargs != 1) fail(1, "line %u: %s needs one probability", line, name);
      if (nargs > 1) fail(1, "line %u: %s takes at most one argument", line, name);
      o.arg = nargs ? args[0] : 0.0;
      if (!(o.arg >= 0.0 && o.arg <= 1.0)) fail(1, "line %u: probability out of range", line);
    }
    if (o.kind == K_OBS) {
      if (nargs != 1 || args[0] < 0 || args[0] != floor(args[0]))
        fail(1, "line %u: OBSERVABLE_INCLUDE needs an index", line);
      o.arg = args[0];
    }
    if ((o.kind == K_CX || o.kind == K_DEP2) && (nt % 2))
      fail(1, "line %u: %s needs target pairs", line, name);
    if (o.kind == K_TICK && nt) fail(1, "line %u: TICK takes no targets", line);
    if (k == 0) {
      o.t0 = t0;
      o.nt = nt;
    } else { /* second half of MR reuses the same targets */
      o.t0 = (uint32_t)out->ntg;
      for (uint32_t j = 0; j < nt; ++j) ol_push_tg(out, out->tg[t0 + j]);
      o.nt = nt;
    }
    ol_push_op(out, o);
  }
}

static void parse_block(lex_t *lx, oplist_t *out, int depth, char **buf,
                        size_t *cap) {
  while (next_line(lx, buf, cap)) {
    if ((*buf)[0] == '}') {
      if (depth == 0) fail(1, "line %u: unmatched '}'", lx->line);
      return;
    }
    parse_instruction(lx, out, *buf, depth, buf, cap);
  }
  if (depth) fail(1, "unterminated REPEAT block");
}

static void circ_parse(const char *text, circ_t *C) {
  memset(C, 0, sizeof *C);
  lex_t lx = {text, 0, strlen(text), 0};
  char *buf = NULL;
  size_t cap = 0;
  parse_block(&lx, &C->L, 0, &buf, &cap);
  free(buf);
  /* stats (stim's compute_stats, used at Circuit.hpp:68-79) */
  int64_t maxq = -1;
  uint32_t meas = 0;
  for (size_t i = 0; i < C->L.nops; ++i) {
    op_t *o = &C->L.ops[i];
    switch (o->kind) {
    case K_TICK: C->nticks++; break;
    case K_DET: C->ndet++; break;
    case K_OBS:
      if ((uint32_t)o->arg + 1 > C->nobs) C->nobs = (uint32_t)o->arg + 1;
      break;
    default: break;
    }
    if (o->kind == K_DET || o->kind == K_OBS) {
      for (uint32_t k = 0; k < o->nt; ++k)
        if ((uint32_t)(-C->L.tg[o->t0 + k]) > meas)
          fail(1, "record target rec[%d] refers before the first measurement", C->L.tg[o->t0 + k]);
    } else if (o->kind != K_TICK) {
      for (uint32_t k = 0; k < o->nt; ++k)
        if (C->L.tg[o->t0 + k] > maxq) maxq = C->L.tg[o->t0 + k];
      if (o->kind == K_M) meas += o->nt;
    }
  }
  C->nq = (uint32_t)(maxq + 1);
  C->nmeas = meas;
}

/* ------------------------------------------------------------------ */
/* noise splits and merge (Mixer.hpp:164, Mixer.hpp:184, STEPG.hpp:128-131)*/
/* ------------------------------------------------------------------ */

double gpo_split_dep1(double p) { return 0.5 - 0.5 * sqrt(1 - (4 * p) / 3); }
double gpo_split_dep2(double p) { return 0.5 - 0.5 * pow(1 - (16 * p) / 15, 0.125); }
double gpo_xor_merge(double p_old, double p_new) {
  return (1 - p_new) * p_old + (1 - p_old) * p_new;
}

/* FNV-1a 64 over n 32-bit slots, padded with 0xFFFFFFFF to pad_to slots
 * (Array.hpp:52-65: offset basis 0xcbf29ce484222325, prime 0x100000001b3,
 * one XOR+multiply per 32-bit slot). */
uint64_t gpo_fnv1a64_slots(const uint32_t *vals, uint32_t n, uint32_t pad_to) {
  uint64_t h = 0xcbf29ce484222325ull;
  const uint64_t prime = 0x100000001b3ull;
  for (uint32_t i = 0; i < n; ++i) {
    h ^= (uint64_t)vals[i];
    h *= prime;
  }
  for (uint32_t i = n; i < pad_to; ++i) {
    h ^= (uint64_t)NONE32;
    h *= prime;
  }
  return h;
}

/* ------------------------------------------------------------------ */
/* STEPG restatement (STEPG.hpp:86-131, Graph.hpp:132-182, Mixer.hpp)    */
/* ------------------------------------------------------------------ */

struct gpo_stepg {
  uint32_t nl, nq, k, npl, nn, M, D, O, W;
  uint64_t *adj;   /* nn: child0 low 32 bits, child1 high 32 bits */
  double *probs;   /* nn */
  uint32_t *keys;  /* M */
  uint64_t *vals;  /* M x W row-major */
  /* per-component bookkeeping for the gradient oracle */
  uint32_t ncomp, capcomp;
  uint32_t *comp_node, *comp_loc;
  uint8_t *comp_kind; /* 0 direct, 1 dep1 split, 2 dep2 split */
  uint32_t nloc;
};

enum { S_X = 0, S_Z = 1, S_XXZZ = 2, S_Y = 3, S_XZZX = 4, S_XYYX = 5, S_YY = 6, S_YZZY = 7 };

static uint32_t nodes_per_qubit(int level) {
  switch (level) {
  case 0: return 3;
  case 1: return 5;
  case 2: return 8;
  }
  fail(1, "Invalid correlation level %d", level);
  return 0;
}

static uint32_t nid(const struct gpo_stepg *g, uint32_t q, uint32_t slot, uint32_t l) {
  return l * g->npl + q * g->k + slot;
}

static void add_flow(struct gpo_stepg *g, uint32_t sq, uint32_t ss, uint32_t sl,
                     uint32_t tq, uint32_t ts, uint32_t tl) {
  if (tl <= sl) fail(2, "STEPG: flows must move forward in time.");
  if (sl >= g->nl || tl >= g->nl) fail(2, "STEPG: graph.addEdge failed with status 3.");
  uint32_t s = nid(g, sq, ss, sl), t = nid(g, tq, ts, tl);
  uint64_t a = g->adj[s];
  uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
  if (lo == t || hi == t) fail(2, "STEPG: graph.addEdge failed with status 1.");
  if (lo == NONE32) g->adj[s] = (a & 0xFFFFFFFF00000000ull) | t;
  else if (hi == NONE32) g->adj[s] = (a & 0xFFFFFFFFull) | ((uint64_t)t << 32);
  else fail(2, "STEPG: graph.addEdge failed with status 2.");
}

static void remove_flow(struct gpo_stepg *g, uint32_t sq, uint32_t ss, uint32_t sl,
                        uint32_t tq, uint32_t ts, uint32_t tl) {
  if (tl <= sl) fail(2, "STEPG: flows must move forward in time.");
  if (sl >= g->nl || tl >= g->nl) fail(2, "STEPG: graph.removeEdge failed with status 3.");
  uint32_t s = nid(g, sq, ss, sl), t = nid(g, tq, ts, tl);
  uint64_t a = g->adj[s];
  if ((uint32_t)a == t) g->adj[s] = a | 0xFFFFFFFFull;
  else if ((uint32_t)(a >> 32) == t) g->adj[s] = a | 0xFFFFFFFF00000000ull;
  else fail(2, "STEPG: graph.removeEdge failed with status 4.");
}

/* Adds aux -> (c0, c1) unless aux already has exactly these children (the
 * same component from another channel in the layer: merge only). */
static void wire_aux(struct gpo_stepg *g, uint32_t aq, uint32_t as, uint32_t al, uint32_t q0,
                     uint32_t s0, uint32_t l0, uint32_t q1, uint32_t s1, uint32_t l1) {
  uint32_t a = nid(g, aq, as, al), c0 = nid(g, q0, s0, l0), c1 = nid(g, q1, s1, l1);
  if ((uint32_t)g->adj[a] == c0 && (uint32_t)(g->adj[a] >> 32) == c1) return;
  add_flow(g, aq, as, al, q0, s0, l0);
  add_flow(g, aq, as, al, q1, s1, l1);
}

static void add_comp(struct gpo_stepg *g, uint32_t node, uint32_t loc, uint8_t kind) {
  if (g->ncomp == g->capcomp) {
    g->capcomp = g->capcomp ? 2 * g->capcomp : 4096;
    g->comp_node = xrealloc(g->comp_node, g->capcomp * sizeof(uint32_t));
    g->comp_loc = xrealloc(g->comp_loc, g->capcomp * sizeof(uint32_t));
    g->comp_kind = xrealloc(g->comp_kind, g->capcomp);
  }
  g->comp_node[g->ncomp] = node;
  g->comp_loc[g->ncomp] = loc;
  g->comp_kind[g->ncomp] = kind;
  g->ncomp++;
}

static void merge_at(struct gpo_stepg *g, uint32_t node, double p, uint32_t loc, uint8_t kind) {
  g->probs[node] = gpo_xor_merge(g->probs[node], p);
  add_comp(g, node, loc, kind);
}

/* Single-qubit Pauli components (Mixer.hpp:209-248). */
static void err_x(struct gpo_stepg *g, int lv, uint32_t q, uint32_t l, double p, uint32_t loc, uint8_t kd) {
  (void)lv;
  merge_at(g, nid(g, q, S_X, l), p, loc, kd);
}
static void err_z(struct gpo_stepg *g, int lv, uint32_t q, uint32_t l, double p, uint32_t loc, uint8_t kd) {
  (void)lv;
  merge_at(g, nid(g, q, S_Z, l), p, loc, kd);
}
static void err_y(struct gpo_stepg *g, int lv, uint32_t q, uint32_t l, double p, uint32_t loc, uint8_t kd) {
  if (lv < 1) return; /* L0 drops Y components */
  if (l < 1) fail(2, "Mixer: cannot apply Y error in layer l < 1");
  wire_aux(g, q, S_Y, l - 1, q, S_X, l, q, S_Z, l);
  merge_at(g, nid(g, q, S_Y, l - 1), p, loc, kd);
}

/* Two-qubit correlated components: an auxiliary node one (or two) layers
 * earlier whose out-flows reproduce the Pauli pair (Mixer.hpp:305-533). */
static void err_pair(struct gpo_stepg *g, int lv, int pa, int pb, uint32_t q0, uint32_t q1,
                     uint32_t l, double p, uint32_t loc) {
  /* pa, pb in {0:I, 1:X, 2:Y, 3:Z} */
  const uint8_t kd = 2;
  if (pa == 0) {
    if (pb == 1) err_x(g, lv, q1, l, p, loc, kd);
    else if (pb == 2) err_y(g, lv, q1, l, p, loc, kd);
    else err_z(g, lv, q1, l, p, loc, kd);
    return;
  }
  if (pb == 0) {
    if (pa == 1) err_x(g, lv, q0, l, p, loc, kd);
    else if (pa == 2) err_y(g, lv, q0, l, p, loc, kd);
    else err_z(g, lv, q0, l, p, loc, kd);
    return;
  }
  if (pa == 1 && pb == 1) { /* XX: q0 slot 2, layer l-1 */
    if (l < 1) fail(2, "Mixer: cannot apply XX error in layer l < 1");
    wire_aux(g, q0, S_XXZZ, l - 1, q0, S_X, l, q1, S_X, l);
    merge_at(g, nid(g, q0, S_XXZZ, l - 1), p, loc, kd);
  } else if (pa == 3 && pb == 3) { /* ZZ: q1 slot 2, layer l-1 */
    if (l < 1) fail(2, "Mixer: cannot apply ZZ error in layer l < 1");
    wire_aux(g, q1, S_XXZZ, l - 1, q0, S_Z, l, q1, S_Z, l);
    merge_at(g, nid(g, q1, S_XXZZ, l - 1), p, loc, kd);
  } else if (pa == 1 && pb == 3) { /* XZ: q0 slot 4 (L1+) */
    if (lv < 1) return;
    if (l < 1) fail(2, "Mixer: cannot apply XZ error in layer l < 1");
    wire_aux(g, q0, S_XZZX, l - 1, q0, S_X, l, q1, S_Z, l);
    merge_at(g, nid(g, q0, S_XZZX, l - 1), p, loc, kd);
  } else if (pa == 3 && pb == 1) { /* ZX: q1 slot 4 (L1+) */
    if (lv < 1) return;
    if (l < 1) fail(2, "Mixer: cannot apply ZX error in layer l < 1");
    wire_aux(g, q1, S_XZZX, l - 1, q0, S_Z, l, q1, S_X, l);
    merge_at(g, nid(g, q1, S_XZZX, l - 1), p, loc, kd);
  } else {
    if (lv < 2) return; /* pairs involving Y need L2 */
    if (l < 2) {
      static const char *nm[4] = {"I", "X", "Y", "Z"};
      fail(2, "Mixer: cannot apply %s%s error in layer l < 2", nm[pa], nm[pb]);
    }
    if (pa == 1 && pb == 2) { /* XY: q0 slot 5 -> X0(l), Y1(l-1) */
      wire_aux(g, q0, S_XYYX, l - 2, q0, S_X, l, q1, S_Y, l - 1);
      merge_at(g, nid(g, q0, S_XYYX, l - 2), p, loc, kd);
    } else if (pa == 2 && pb == 1) { /* YX: q1 slot 5 -> Y0(l-1), X1(l) */
      wire_aux(g, q1, S_XYYX, l - 2, q0, S_Y, l - 1, q1, S_X, l);
      merge_at(g, nid(g, q1, S_XYYX, l - 2), p, loc, kd);
    } else if (pa == 2 && pb == 2) { /* YY: q0 slot 6 -> Y0(l-1), Y1(l-1) */
      wire_aux(g, q0, S_YY, l - 2, q0, S_Y, l - 1, q1, S_Y, l - 1);
      merge_at(g, nid(g, q0, S_YY, l - 2), p, loc, kd);
    } else if (pa == 2 && pb == 3) { /* YZ: q1 slot 7 -> Y0(l-1), Z1(l) */
      wire_aux(g, q1, S_YZZY, l - 2, q0, S_Y, l - 1, q1, S_Z, l);
      merge_at(g, nid(g, q1, S_YZZY, l - 2), p, loc, kd);
    } else { /* ZY: q0 slot 7 -> Z0(l), Y1(l-1) */
      wire_aux(g, q0, S_YZZY, l - 2, q0, S_Z, l, q1, S_Y, l - 1);
      merge_at(g, nid(g, q0, S_YZZY, l - 2), p, loc, kd);
    }
  }
}

static void stepg_free_inner(struct gpo_stepg *g) {
  free(g->adj);
  free(g->probs);
  free(g->keys);
  free(g->vals);
  free(g->comp_node);
  free(g->comp_loc);
  free(g->comp_kind);
}

void gpo_stepg_free(gpo_stepg *g) {
  if (!g) return;
  stepg_free_inner(g);
  free(g);
}

/* Builds the STEPG (Circuit.hpp:471-478).  theta, if given, overrides the
 * probability of each fault location (in circuit order). */
static void stepg_build(const circ_t *C, int level, const double *theta, uint32_t ntheta,
                        struct gpo_stepg *g) {
  memset(g, 0, sizeof *g);
  g->k = nodes_per_qubit(level);
  g->nl = C->nticks + 1;
  g->nq = C->nq;
  g->npl = g->nq * g->k;
  uint64_t nn = (uint64_t)g->nl * g->npl;
  if (nn > 0x7FFFFFFFull) fail(2, "circuit too large for 32-bit node indices");
  g->nn = (uint32_t)nn;
  g->M = C->nmeas;
  g->D = C->ndet;
  g->O = C->nobs;
  g->W = (g->D + g->O + 63) / 64;
  g->adj = xmalloc((size_t)g->nn * 8);
  memset(g->adj, 0xFF, (size_t)g->nn * 8);
  g->probs = xcalloc(g->nn, 8);
  g->keys = xcalloc(g->M, 4);
  g->vals = xcalloc((size_t)g->M * (g->W ? g->W : 1), 8);

  /* persistence flows (Mixer.hpp:97-108) */
  for (uint32_t l = 0; l + 1 < g->nl; ++l)
    for (uint32_t q = 0; q < g->nq; ++q) {
      add_flow(g, q, S_X, l, q, S_X, l + 1);
      add_flow(g, q, S_Z, l, q, S_Z, l + 1);
    }

  uint32_t layer = 0, meas = 0, det = 0, loc = 0;
  const int32_t *tg = C->L.tg;
  for (size_t i = 0; i < C->L.nops; ++i) {
    const op_t *o = &C->L.ops[i];
    const int32_t *t = tg + o->t0;
    switch (o->kind) {
    case K_TICK: layer++; break;
    case K_R: /* Mixer.hpp:114-120; no later layer after the last: no-op */
      if (layer + 1 >= g->nl) break;
      for (uint32_t j = 0; j < o->nt; ++j) {
        remove_flow(g, t[j], S_X, layer, t[j], S_X, layer + 1);
        remove_flow(g, t[j], S_Z, layer, t[j], S_Z, layer + 1);
      }
      break;
    case K_CX: /* Mixer.hpp:127-136 */
      if (layer + 1 >= g->nl) break;
      for (uint32_t j = 0; j < o->nt; j += 2) {
        uint32_t c = t[j], tt = t[j + 1];
        add_flow(g, c, S_X, layer, tt, S_X, layer + 1);
        add_flow(g, tt, S_Z, layer, c, S_Z, layer + 1);
      }
      break;
    case K_H: /* Mixer.hpp:142-150 */
      if (layer + 1 >= g->nl) break;
      for (uint32_t j = 0; j < o->nt; ++j) {
        uint32_t q = t[j];
        remove_flow(g, q, S_X, layer, q, S_X, layer + 1);
        remove_flow(g, q, S_Z, layer, q, S_Z, layer + 1);
        add_flow(g, q, S_X, layer, q, S_Z, layer + 1);
        add_flow(g, q, S_Z, layer, q, S_X, layer + 1);
      }
      break;
    case K_M: /* Circuit.hpp:316-319: X error on the measured qubit + key */
      for (uint32_t j = 0; j < o->nt; ++j) {
        double p = o->arg;
        if (theta) {
          if (loc >= ntheta) fail(1, "theta has too few entries");
          p = theta[loc];
        }
        err_x(g, level, t[j], layer, p, loc, 0);
        g->keys[meas++] = nid(g, t[j], S_X, layer);
        loc++;
      }
      break;
    case K_XERR: /* Mixer.hpp:209-214 */
      for (uint32_t j = 0; j < o->nt; ++j) {
        double p = theta ? (loc < ntheta ? theta[loc] : (fail(1, "theta has too few entries"), 0)) : o->arg;
        err_x(g, level, t[j], layer, p, loc, 0);
        loc++;
      }
      break;
    case K_DEP1: /* Mixer.hpp:158-169: X, Z, Y each with the split */
      for (uint32_t j = 0; j < o->nt; ++j) {
        double p = theta ? (loc < ntheta ? theta[loc] : (fail(1, "theta has too few entries"), 0)) : o->arg;
        if (p > 0.75) fail(2, "Mixer: invalid DEP1 p > 0.75.");
        double s = gpo_split_dep1(p);
        err_x(g, level, t[j], layer, s, loc, 1);
        err_z(g, level, t[j], layer, s, loc, 1);
        err_y(g, level, t[j], layer, s, loc, 1);
        loc++;
      }
      break;
    case K_DEP2: /* Mixer.hpp:178-201, fixed component order */
      for (uint32_t j = 0; j < o->nt; j += 2) {
        double p = theta ? (loc < ntheta ? theta[loc] : (fail(1, "theta has too few entries"), 0)) : o->arg;
        if (p > 0.9375) fail(2, "Mixer: invalid DEP2 p > 0.9375.");
        double s = gpo_split_dep2(p);
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b)
            if (a || b) err_pair(g, level, a, b, t[j], t[j + 1], layer, s, loc);
        loc++;
      }
      break;
    case K_DET:
    case K_OBS: {
      uint32_t bit = o->kind == K_DET ? det++ : g->D + (uint32_t)o->arg;
      for (uint32_t j = 0; j < o->nt; ++j) {
        uint32_t row = meas + (uint32_t)t[j]; /* t[j] < 0 */
        g->vals[(size_t)row * g->W + bit / 64] ^= 1ull << (bit % 64);
      }
      break;
    }
    }
  }
  g->nloc = loc;
}

int gpo_build_stepg(const char *text, int level, gpo_stepg **out) {
  jmp_buf jb;
  jmp_buf *saved = g_jmp;
  g_jmp = &jb;
  g_err[0] = 0;
  circ_t C;
  memset(&C, 0, sizeof C);
  struct gpo_stepg *g = NULL;
  if (setjmp(jb)) {
    ol_free(&C.L);
    if (g) gpo_stepg_free(g);
    g_jmp = saved;
    return g_code;
  }
  circ_parse(text, &C);
  g = xcalloc(1, sizeof *g);
  stepg_build(&C, level, NULL, 0, g);
  ol_free(&C.L);
  *out = g;
  g_jmp = saved;
  return 0;
}

void gpo_stepg_dims(const gpo_stepg *g, uint32_t *d) {
  d[0] = g->nl; d[1] = g->nq; d[2] = g->k; d[3] = g->nn;
  d[4] = g->M; d[5] = g->D; d[6] = g->O; d[7] = g->W;
}

void gpo_stepg_copy(const gpo_stepg *g, uint64_t *adj, double *probs, uint32_t *keys,
                    uint64_t *vals) {
  if (adj) memcpy(adj, g->adj, (size_t)g->nn * 8);
  if (probs) memcpy(probs, g->probs, (size_t)g->nn * 8);
  if (keys) memcpy(keys, g->keys, (size_t)g->M * 4);
  if (vals) memcpy(vals, g->vals, (size_t)g->M * g->W * 8);
}

/* Backward sweep (CUDA.hpp:37-51 + CUDAHandlers.hpp:46-61), seed XOR
 * children, layer by layer from the last layer down. */
static void sweep(const struct gpo_stepg *g, uint64_t *S) {
  const uint32_t W = g->W;
  memset(S, 0, (size_t)g->nn * W * 8);
  for (uint32_t m = 0; m < g->M; ++m)
    for (uint32_t w = 0; w < W; ++w) S[(size_t)g->keys[m] * W + w] ^= g->vals[(size_t)m * W + w];
  for (int64_t l = (int64_t)g->nl - 1; l >= 0; --l) {
    for (uint32_t j = 0; j < g->npl; ++j) {
      uint32_t n = (uint32_t)l * g->npl + j;
      uint32_t c0 = (uint32_t)g->adj[n], c1 = (uint32_t)(g->adj[n] >> 32);
      for (uint32_t w = 0; w < W; ++w) {
        uint64_t s = S[(size_t)n * W + w];
        if (c0 != NONE32) s ^= S[(size_t)c0 * W + w];
        if (c1 != NONE32) s ^= S[(size_t)c1 * W + w];
        S[(size_t)n * W + w] = s;
      }
    }
  }
}

void gpo_stepg_sweep(const gpo_stepg *g, uint64_t *S) { sweep(g, S); }

/* ------------------------------------------------------------------ */
/* DEM container                                                        */
/* ------------------------------------------------------------------ */

struct gpo_dem {
  uint32_t n, D, O;
  uint64_t nt;
  double *p;
  uint32_t *off;
  uint32_t *tg;
  uint64_t *hash;
};

void gpo_dem_free(gpo_dem *d) {
  if (!d) return;
  free(d->p);
  free(d->off);
  free(d->tg);
  free(d->hash);
  free(d);
}
uint32_t gpo_dem_num_errors(const gpo_dem *d) { return d->n; }
uint64_t gpo_dem_num_targets(const gpo_dem *d) { return d->nt; }
uint32_t gpo_dem_num_detectors(const gpo_dem *d) { return d->D; }
uint32_t gpo_dem_num_observables(const gpo_dem *d) { return d->O; }
void gpo_dem_copy(const gpo_dem *d, double *p, uint32_t *off, uint32_t *tg) {
  if (p) memcpy(p, d->p, (size_t)d->n * 8);
  if (off) memcpy(off, d->off, (size_t)(d->n + 1) * 4);
  if (tg) memcpy(tg, d->tg, (size_t)d->nt * 4);
}
void gpo_dem_hashes(const gpo_dem *d, uint64_t *h) {
  if (d->hash) memcpy(h, d->hash, (size_t)d->n * 8);
  else memset(h, 0, (size_t)d->n * 8);
}

/* ------------------------------------------------------------------ */
/* STEPG compiler: classes, hash, stable sort, fold                      */
/* ------------------------------------------------------------------ */

typedef struct {
  uint64_t h;
  uint32_t node;
} hn_t;

static int cmp_hn(const void *a, const void *b) {
  const hn_t *x = a, *y = b;
  if (x->h != y->h) return x->h < y->h ? -1 : 1;
  return x->node < y->node ? -1 : (x->node > y->node);
}

static uint32_t class_of(const uint64_t *row, uint32_t W, uint32_t *out) {
  uint32_t n = 0;
  for (uint32_t w = 0; w < W; ++w) {
    uint64_t v = row[w];
    while (v) {
      out[n++] = w * 64 + (uint32_t)__builtin_ctzll(v);
      v &= v - 1;
    }
  }
  return n;
}

typedef struct {
  uint32_t C;
  uint32_t *cls_of_node; /* nn, NONE32 if dropped */
  hn_t *sorted;          /* N' sorted (hash,node) */
  uint32_t nprime;
  uint32_t *run_start;   /* C+1 */
} grouping_t;

static void group_nodes(const struct gpo_stepg *g, const double *probs, const uint64_t *S,
                        grouping_t *G, gpo_dem *d) {
  const uint32_t W = g->W;
  uint32_t *buf = xmalloc((size_t)(W * 64 + 1) * 4);
  hn_t *hn = xmalloc((size_t)g->nn * sizeof(hn_t));
  uint32_t np = 0;
  for (uint32_t n = 0; n < g->nn; ++n) {
    if (!(probs[n] > 0)) continue;
    uint32_t c = class_of(S + (size_t)n * W, W, buf);
    if (!c) continue;
    hn[np].h = gpo_fnv1a64_slots(buf, c, c <= 32 ? 32 : c);
    hn[np].node = n;
    np++;
  }
  qsort(hn, np, sizeof(hn_t), cmp_hn);
  uint32_t C = 0;
  for (uint32_t i = 0; i < np; ++i)
    if (i == 0 || hn[i].h != hn[i - 1].h) C++;
  G->C = C;
  G->sorted = hn;
  G->nprime = np;
  G->run_start = xmalloc((size_t)(C + 1) * 4);
  G->cls_of_node = xmalloc((size_t)g->nn * 4);
  memset(G->cls_of_node, 0xFF, (size_t)g->nn * 4);
  uint32_t r = 0;
  for (uint32_t i = 0; i < np; ++i) {
    if (i == 0 || hn[i].h != hn[i - 1].h) G->run_start[r++] = i;
    G->cls_of_node[hn[i].node] = r - 1;
  }
  G->run_start[C] = np;

  /* DEM: one error per run, rep = first (smallest) node, sequential fold */
  d->n = C;
  d->D = g->D;
  d->O = g->O;
  d->p = xmalloc((size_t)C * 8 + 8);
  d->off = xmalloc((size_t)(C + 1) * 4);
  d->hash = xmalloc((size_t)C * 8 + 8);
  uint64_t nt = 0;
  for (uint32_t c = 0; c < C; ++c) nt += class_of(S + (size_t)hn[G->run_start[c]].node * W, W, buf);
  d->tg = xmalloc(nt * 4 + 4);
  d->nt = nt;
  uint64_t pos = 0;
  for (uint32_t c = 0; c < C; ++c) {
    uint32_t a = G->run_start[c], b = G->run_start[c + 1];
    double p = probs[hn[a].node];
    for (uint32_t i = a + 1; i < b; ++i) {
      /* reducer of CUDAHandlers.hpp:133-138: p = (1-p1) p0 + (1-p0) p1 */
      double p1 = probs[hn[i].node];
      p = (1 - p1) * p + (1 - p) * p1;
    }
    /* collision check: every member must carry the rep's exact class */
    const uint64_t *rep = S + (size_t)hn[a].node * W;
    for (uint32_t i = a + 1; i < b; ++i)
      if (memcmp(rep, S + (size_t)hn[i].node * W, (size_t)W * 8))
        fail(2, "64-bit class hash collision between nodes %u and %u", hn[a].node, hn[i].node);
    d->p[c] = p;
    d->hash[c] = hn[a].h;
    d->off[c] = (uint32_t)pos;
    pos += class_of(rep, W, d->tg + pos);
  }
  d->off[C] = (uint32_t)pos;
  free(buf);
}

static void grouping_free(grouping_t *G) {
  free(G->cls_of_node);
  free(G->sorted);
  free(G->run_start);
}

int gpo_compile_stepg(const char *text, int level, gpo_dem **out) {
  return gpo_compile_stepg_theta(text, level, NULL, 0, out);
}

int gpo_compile_stepg_theta(const char *text, int level, const double *theta,
                            uint32_t num_theta, gpo_dem **out) {
  jmp_buf jb;
  jmp_buf *saved = g_jmp;
  g_jmp = &jb;
  g_err[0] = 0;
  circ_t C;
  memset(&C, 0, sizeof C);
  struct gpo_stepg g;
  memset(&g, 0, sizeof g);
  uint64_t *S = NULL;
  gpo_dem *d = NULL;
  grouping_t G;
  memset(&G, 0, sizeof G);
  if (setjmp(jb)) {
    ol_free(&C.L);
    stepg_free_inner(&g);
    free(S);
    grouping_free(&G);
    gpo_dem_free(d);
    g_jmp = saved;
    return g_code;
  }
  circ_parse(text, &C);
  stepg_build(&C, level, theta, num_theta, &g);
  S = xmalloc((size_t)g.nn * (g.W ? g.W : 1) * 8);
  sweep(&g, S);
  d = xcalloc(1, sizeof *d);
  group_nodes(&g, g.probs, S, &G, d);
  grouping_free(&G);
  free(S);
  stepg_free_inner(&g);
  ol_free(&C.L);
  *out = d;
  g_jmp = saved;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Gradient oracle: L = sum_e w_e logit(p_e(theta))                      */
/* ------------------------------------------------------------------ */

int gpo_location_count(const char *text, uint32_t *count) {
  jmp_buf jb;
  jmp_buf *saved = g_jmp;
  g_jmp = &jb;
  g_err[0] = 0;
  circ_t C;
  memset(&C, 0, sizeof C);
  if (setjmp(jb)) {
    ol_free(&C.L);
    g_jmp = saved;
    return g_code;
  }
  circ_parse(text, &C);
  uint32_t n = 0;
  for (size_t i = 0; i < C.L.nops; ++i) {
    const op_t *o = &C.L.ops[i];
    if (o->kind == K_M || o->kind == K_XERR || o->kind == K_DEP1) n += o->nt;
    else if (o->kind == K_DEP2) n += o->nt / 2;
  }
  ol_free(&C.L);
  *count = n;
  g_jmp = saved;
  return 0;
}

int gpo_logit_loss_grad(const char *text, int level, const double *theta, uint32_t nloc_in,
                        const double *w, uint32_t nw, double *loss, double *grad,
                        uint32_t nloc_out) {
  jmp_buf jb;
  jmp_buf *saved = g_jmp;
  g_jmp = &jb;
  g_err[0] = 0;
  circ_t C;
  memset(&C, 0, sizeof C);
  struct gpo_stepg g;
  memset(&g, 0, sizeof g);
  uint64_t *S = NULL;
  gpo_dem *d = NULL;
  grouping_t G;
  memset(&G, 0, sizeof G);
  double *r = NULL, *dr = NULL;
  uint32_t *nstart = NULL, *order = NULL;
  if (setjmp(jb)) {
    ol_free(&C.L);
    stepg_free_inner(&g);
    free(S);
    grouping_free(&G);
    gpo_dem_free(d);
    free(r); free(dr); free(nstart); free(order);
    g_jmp = saved;
    return g_code;
  }
  circ_parse(text, &C);
  stepg_build(&C, level, theta, nloc_in, &g);
  if (g.nloc != nloc_out) fail(1, "gradient has %u entries, circuit has %u locations", nloc_out, g.nloc);
  S = xmalloc((size_t)g.nn * (g.W ? g.W : 1) * 8);
  sweep(&g, S);
  d = xcalloc(1, sizeof *d);
  group_nodes(&g, g.probs, S, &G, d);
  if (nw != G.C) fail(1, "weights have %u entries, DEM has %u errors", nw, G.C);

  /* loss and dL/dp_e */
  double L = 0;
  for (uint32_t e = 0; e < G.C; ++e) L += w[e] * log(d->p[e] / (1 - d->p[e]));
  *loss = L;

  /* component probabilities and derivatives w.r.t. their location rate */
  r = xmalloc((size_t)g.ncomp * 8 + 8);
  dr = xmalloc((size_t)g.ncomp * 8 + 8);
  for (uint32_t c = 0; c < g.ncomp; ++c) {
    double th = theta ? theta[g.comp_loc[c]] : NAN;
    if (!theta) fail(1, "theta required");
    if (g.comp_kind[c] == 0) { r[c] = th; dr[c] = 1; }
    else if (g.comp_kind[c] == 1) { r[c] = gpo_split_dep1(th); dr[c] = (1.0 / 3.0) / sqrt(1 - 4 * th / 3); }
    else { r[c] = gpo_split_dep2(th); dr[c] = (1.0 / 15.0) * pow(1 - 16 * th / 15, -0.875); }
  }
  /* components grouped by node (counting sort, stable in component order) */
  nstart = xcalloc((size_t)g.nn + 1, 4);
  order = xmalloc((size_t)g.ncomp * 4 + 4);
  for (uint32_t c = 0; c < g.ncomp; ++c) nstart[g.comp_node[c] + 1]++;
  for (uint32_t n = 0; n < g.nn; ++n) nstart[n + 1] += nstart[n];
  {
    uint32_t *cur = xmalloc((size_t)g.nn * 4 + 4);
    memcpy(cur, nstart, (size_t)g.nn * 4);
    for (uint32_t c = 0; c < g.ncomp; ++c) order[cur[g.comp_node[c]]++] = c;
    free(cur);
  }
  for (uint32_t k = 0; k < nloc_out; ++k) grad[k] = 0;
  for (uint32_t e = 0; e < G.C; ++e) {
    double pe = d->p[e];
    double gp = w[e] / (pe * (1 - pe));
    for (uint32_t i = G.run_start[e]; i < G.run_start[e + 1]; ++i) {
      uint32_t node = G.sorted[i].node;
      /* dp_e/dq_node = prod over the other nodes of the class of (1-2q) */
      double others = 1;
      for (uint32_t j = G.run_start[e]; j < G.run_start[e + 1]; ++j)
        if (j != i) others *= 1 - 2 * g.probs[G.sorted[j].node];
      for (uint32_t a = nstart[node]; a < nstart[node + 1]; ++a) {
        uint32_t c = order[a];
        double oc = 1;
        for (uint32_t b = nstart[node]; b < nstart[node + 1]; ++b)
          if (b != a) oc *= 1 - 2 * r[order[b]];
        grad[g.comp_loc[c]] += gp * others * oc * dr[c];
      }
    }
  }
  free(r); free(dr); free(nstart); free(order);
  r = dr = NULL; nstart = order = NULL;
  grouping_free(&G);
  gpo_dem_free(d);
  free(S);
  stepg_free_inner(&g);
  ol_free(&C.L);
  g_jmp = saved;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Forward Pauli-frame compiler (stim semantics, independent of STEPG)   */
/* ------------------------------------------------------------------ */

typedef struct {
  uint32_t op;  /* op index where the component is injected */
  uint32_t q0, q1;
  uint8_t pa, pb; /* Pauli on q0, q1: 0 I, 1 X, 2 Y, 3 Z; pa=4 means record flip */
  uint32_t meas;  /* for record flips */
  double p;
} comp_t;

typedef struct {
  double p;
  uint32_t off, n;
} sym_t;

static const uint32_t *g_sym_pool;
static int cmp_sym(const void *a, const void *b) {
  const sym_t *x = a, *y = b;
  uint32_t n = x->n < y->n ? x->n : y->n;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t u = g_sym_pool[x->off + i], v = g_sym_pool[y->off + i];
    if (u != v) return u < v ? -1 : 1;
  }
  if (x->n != y->n) return x->n < y->n ? -1 : 1;
  return x->off < y->off ? -1 : (x->off > y->off); /* stable */
}

int gpo_compile_frames(const char *text, gpo_dem **out) {
  jmp_buf jb;
  jmp_buf *saved = g_jmp;
  g_jmp = &jb;
  g_err[0] = 0;
  circ_t C;
  memset(&C, 0, sizeof C);
  comp_t *cp = NULL;
  sym_t *sy = NULL;
  uint32_t *pool = NULL;
  uint64_t *x = NULL, *z = NULL, *rec = NULL, *det = NULL, *obs = NULL;
  uint32_t *meas_base = NULL, *det_base = NULL;
  gpo_dem *d = NULL;
  if (setjmp(jb)) {
    ol_free(&C.L);
    free(cp); free(sy); free(pool); free(x); free(z); free(rec); free(det); free(obs);
    free(meas_base); free(det_base);
    gpo_dem_free(d);
    g_jmp = saved;
    return g_code;
  }
  circ_parse(text, &C);
  const uint32_t D = C.ndet, O = C.nobs, Q = C.nq;
  const size_t nops = C.L.nops;
  const int32_t *tg = C.L.tg;

  /* enumerate independent components (stim's channel decomposition) */
  size_t ncp = 0, capcp = 1024;
  cp = xmalloc(capcp * sizeof(comp_t));
  meas_base = xmalloc((nops + 1) * 4);
  det_base = xmalloc((nops + 1) * 4);
  uint32_t m = 0, dd = 0;
  for (size_t i = 0; i < nops; ++i) {
    const op_t *o = &C.L.ops[i];
    const int32_t *t = tg + o->t0;
    meas_base[i] = m;
    det_base[i] = dd;
#define PUSH(c_)                                                      \
  do {                                                                \
    if (ncp == capcp) { capcp *= 2; cp = xrealloc(cp, capcp * sizeof(comp_t)); } \
    cp[ncp++] = (c_);                                                 \
  } while (0)
    if (o->kind == K_XERR) {
      for (uint32_t j = 0; j < o->nt; ++j) PUSH(((comp_t){(uint32_t)i, t[j], 0, 1, 0, 0, o->arg}));
    } else if (o->kind == K_DEP1) {
      if (o->arg > 0.75) fail(2, "invalid DEPOLARIZE1 p > 0.75");
      double s = gpo_split_dep1(o->arg);
      for (uint32_t j = 0; j < o->nt; ++j)
        for (uint8_t a = 1; a <= 3; ++a) PUSH(((comp_t){(uint32_t)i, t[j], 0, a, 0, 0, s}));
    } else if (o->kind == K_DEP2) {
      if (o->arg > 0.9375) fail(2, "invalid DEPOLARIZE2 p > 0.9375");
      double s = gpo_split_dep2(o->arg);
      for (uint32_t j = 0; j < o->nt; j += 2)
        for (uint8_t a = 0; a < 4; ++a)
          for (uint8_t b = 0; b < 4; ++b)
            if (a || b) PUSH(((comp_t){(uint32_t)i, t[j], t[j + 1], a, b, 0, s}));
    } else if (o->kind == K_M) {
      for (uint32_t j = 0; j < o->nt; ++j) PUSH(((comp_t){(uint32_t)i, t[j], 0, 4, 0, m + j, o->arg}));
      m += o->nt;
    } else if (o->kind == K_DET) {
      dd++;
    }
#undef PUSH
  }
  meas_base[nops] = m;
  det_base[nops] = dd;

  x = xcalloc(Q ? Q : 1, 8);
  z = xcalloc(Q ? Q : 1, 8);
  rec = xcalloc(C.nmeas ? C.nmeas : 1, 8);
  det = xcalloc(D ? D : 1, 8);
  obs = xcalloc(O ? O : 1, 8);
  size_t capsy = 1024, nsy = 0, cappool = 4096, npool = 0;
  sy = xmalloc(capsy * sizeof(sym_t));
  pool = xmalloc(cappool * 4);

  size_t c0 = 0;
  while (c0 < ncp) {
    /* skip zero-probability components: they never appear in a DEM here */
    size_t c1 = c0;
    uint32_t lanes = 0;
    size_t idx[64];
    while (c1 < ncp && lanes < 64) {
      if (cp[c1].p > 0) idx[lanes++] = c1;
      c1++;
    }
    if (!lanes) { c0 = c1; continue; }
    uint32_t start = cp[idx[0]].op;
    memset(x, 0, (size_t)Q * 8);
    memset(z, 0, (size_t)Q * 8);
    memset(det, 0, (size_t)D * 8);
    memset(obs, 0, (size_t)O * 8);
    /* records before the start op carry no flips */
    memset(rec, 0, (size_t)C.nmeas * 8);
    uint32_t li = 0;
    for (size_t i = start; i < nops; ++i) {
      const op_t *o = &C.L.ops[i];
      const int32_t *t = tg + o->t0;
      /* inject the components that live at this op (before the op acts) */
      while (li < lanes && cp[idx[li]].op == i) {
        const comp_t *c = &cp[idx[li]];
        uint64_t bit = 1ull << li;
        if (c->pa == 4) {
          rec[c->meas] ^= bit; /* measurement-result flip */
        } else {
          if (c->pa == 1 || c->pa == 2) x[c->q0] ^= bit;
          if (c->pa == 3 || c->pa == 2) z[c->q0] ^= bit;
          if (c->pb == 1 || c->pb == 2) x[c->q1] ^= bit;
          if (c->pb == 3 || c->pb == 2) z[c->q1] ^= bit;
        }
        li++;
      }
      switch (o->kind) {
      case K_H:
        for (uint32_t j = 0; j < o->nt; ++j) {
          uint64_t tmp = x[t[j]];
          x[t[j]] = z[t[j]];
          z[t[j]] = tmp;
        }
        break;
      case K_CX:
        for (uint32_t j = 0; j < o->nt; j += 2) {
          x[t[j + 1]] ^= x[t[j]];
          z[t[j]] ^= z[t[j + 1]];
        }
        break;
      case K_R:
        for (uint32_t j = 0; j < o->nt; ++j) x[t[j]] = z[t[j]] = 0;
        break;
      case K_M:
        for (uint32_t j = 0; j < o->nt; ++j) {
          rec[meas_base[i] + j] ^= x[t[j]];
          z[t[j]] = 0; /* Z on a Z eigenstate is a phase */
        }
        break;
      case K_DET: {
        uint64_t v = 0;
        for (uint32_t j = 0; j < o->nt; ++j) v ^= rec[meas_base[i] + (uint32_t)t[j]];
        det[det_base[i]] = v;
        break;
      }
      case K_OBS: {
        uint64_t v = 0;
        for (uint32_t j = 0; j < o->nt; ++j) v ^= rec[meas_base[i] + (uint32_t)t[j]];
        obs[(uint32_t)o->arg] ^= v;
        break;
      }
      default: break;
      }
    }
    /* read off symptoms */
    uint32_t cnt[64] = {0};
    for (uint32_t k = 0; k < D; ++k)
      for (uint64_t v = det[k]; v; v &= v - 1) cnt[__builtin_ctzll(v)]++;
    for (uint32_t k = 0; k < O; ++k)
      for (uint64_t v = obs[k]; v; v &= v - 1) cnt[__builtin_ctzll(v)]++;
    uint32_t base[64];
    size_t need = 0;
    for (uint32_t l = 0; l < lanes; ++l) { base[l] = (uint32_t)(npool + need); need += cnt[l]; }
    while (npool + need > cappool) { cappool *= 2; pool = xrealloc(pool, cappool * 4); }
    uint32_t fill[64] = {0};
    for (uint32_t k = 0; k < D; ++k)
      for (uint64_t v = det[k]; v; v &= v - 1) {
        uint32_t l = __builtin_ctzll(v);
        pool[base[l] + fill[l]++] = k;
      }
    for (uint32_t k = 0; k < O; ++k)
      for (uint64_t v = obs[k]; v; v &= v - 1) {
        uint32_t l = __builtin_ctzll(v);
        pool[base[l] + fill[l]++] = D + k;
      }
    npool += need;
    for (uint32_t l = 0; l < lanes; ++l) {
      if (!cnt[l]) continue;
      if (nsy == capsy) { capsy *= 2; sy = xrealloc(sy, capsy * sizeof(sym_t)); }
      sy[nsy++] = (sym_t){cp[idx[l]].p, base[l], cnt[l]};
    }
    c0 = c1;
  }
  g_sym_pool = pool;
  qsort(sy, nsy, sizeof(sym_t), cmp_sym);
  /* merge identical symptoms */
  d = xcalloc(1, sizeof *d);
  d->D = D;
  d->O = O;
  uint32_t ne = 0;
  uint64_t nt = 0;
  for (size_t i = 0; i < nsy; ++i) {
    int same = i > 0 && sy[i].n == sy[i - 1].n &&
               !memcmp(pool + sy[i].off, pool + sy[i - 1].off, (size_t)sy[i].n * 4);
    if (!same) { ne++; nt += sy[i].n; }
  }
  d->n = ne;
  d->nt = nt;
  d->p = xmalloc((size_t)ne * 8 + 8);
  d->off = xmalloc((size_t)(ne + 1) * 4);
  d->tg = xmalloc(nt * 4 + 4);
  uint32_t e = 0;
  uint64_t pos = 0;
  for (size_t i = 0; i < nsy; ++i) {
    int same = i > 0 && sy[i].n == sy[i - 1].n &&
               !memcmp(pool + sy[i].off, pool + sy[i - 1].off, (size_t)sy[i].n * 4);
    if (!same) {
      d->off[e] = (uint32_t)pos;
      memcpy(d->tg + pos, pool + sy[i].off, (size_t)sy[i].n * 4);
      pos += sy[i].n;
      d->p[e] = sy[i].p;
      e++;
    } else {
      d->p[e - 1] = gpo_xor_merge(d->p[e - 1], sy[i].p);
    }
  }
  d->off[ne] = (uint32_t)pos;
  free(cp); free(sy); free(pool); free(x); free(z); free(rec); free(det); free(obs);
  free(meas_base); free(det_base);
  ol_free(&C.L);
  *out = d;
  g_jmp = saved;
  return 0;
}
