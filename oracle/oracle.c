/*
 * oracle.c -- CPU restatement of Replay's Monte Carlo SEIR step loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See oracle.h for the scope, the
 * cited reference files and the documented divergences.
 *
 * Pinning of each part (what checks this restatement, besides reading):
 *   parse / node index / NCLS build / find_intervals: against the reference's
 *       own tcm.c + intervaldb.c compiled into oracle/_ref (tests/golden), and
 *       against a brute-force overlap scan;
 *   window arithmetic: closed forms (steps per window, last-window divisor);
 *   CSR build: dense N x N brute force on small graphs;
 *   exposure: FP64 product within 1e-6 relative, and exactness for integer
 *       weights (integer sums below 2^24 are exact in any order);
 *   or_expf: glibc expf / long-double exp over the used domain;
 *   Philox4x32-10: Random123 known-answer vectors and torch's PhiloxRNGEngine;
 *   counts: brute force over the state with the cu:530-550 classifier;
 *   the loop: closed-form trajectories (p=0 and p=1 cases, pure timers).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAX_LINE_LEN 100000 /* tcm:20 */

/* ------------------------------------------------------------------------ */
/* node-id -> index map (restates tcm:328-356: index = order of first       */
/* appearance; the reference keys uthash by the "%d" string of the id,      */
/* which is one-to-one with the int value).                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t *keys;
  int32_t *vals;
  uint8_t *used;
  size_t cap;
  size_t count;
} idmap;

static uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

static int idmap_init(idmap *m, size_t hint) {
  size_t cap = 64;
  while (cap < hint * 2) cap <<= 1;
  m->keys = (int64_t *)malloc(cap * sizeof(int64_t));
  m->vals = (int32_t *)malloc(cap * sizeof(int32_t));
  m->used = (uint8_t *)calloc(cap, 1);
  m->cap = cap;
  m->count = 0;
  return (m->keys && m->vals && m->used) ? 0 : -1;
}

static void idmap_free(idmap *m) {
  free(m->keys);
  free(m->vals);
  free(m->used);
}

static int idmap_grow(idmap *m);

/* returns index, creating one in first-appearance order */
static int32_t idmap_get_or_create(idmap *m, int64_t key) {
  if ((m->count + 1) * 2 > m->cap && idmap_grow(m) != 0) return -1;
  size_t mask = m->cap - 1;
  size_t h = (size_t)mix64((uint64_t)key) & mask;
  while (m->used[h]) {
    if (m->keys[h] == key) return m->vals[h];
    h = (h + 1) & mask;
  }
  m->used[h] = 1;
  m->keys[h] = key;
  m->vals[h] = (int32_t)m->count;
  m->count++;
  return m->vals[h];
}

static int idmap_grow(idmap *m) {
  idmap n;
  if (idmap_init(&n, m->cap) != 0) return -1;
  for (size_t i = 0; i < m->cap; i++) {
    if (!m->used[i]) continue;
    size_t mask = n.cap - 1;
    size_t h = (size_t)mix64((uint64_t)m->keys[i]) & mask;
    while (n.used[h]) h = (h + 1) & mask;
    n.used[h] = 1;
    n.keys[h] = m->keys[i];
    n.vals[h] = m->vals[i];
  }
  n.count = m->count;
  idmap_free(m);
  *m = n;
  return 0;
}

/* tcm:80-85 */
int64_t or_pack_pair(int a, int b) {
  uint32_t lo = (uint32_t)(a < b ? a : b);
  uint32_t hi = (uint32_t)(a > b ? a : b);
  return (int64_t)(((uint64_t)lo << 32) | hi);
}

/* tcm:87-90 */
static void unpack_pair(int64_t packed, int *src, int *tgt) {
  *tgt = (int)(packed & 0xFFFFFFFF);
  *src = (int)((packed >> 32) & 0xFFFFFFFF);
}

/* tcm:95-120: sscanf of six integers, then timegm (UTC, fields normalised). */
int64_t or_parse_iso8601(const char *s) {
  if (!s) return -1;
  struct tm t;
  memset(&t, 0, sizeof t);
  int year, month, day, hour, min, sec;
  char extra[10];
  if (sscanf(s, "%d-%d-%dT%d:%d:%d%9s", &year, &month, &day, &hour, &min, &sec,
             extra) < 6)
    return -1;
  t.tm_year = year - 1900;
  t.tm_mon = month - 1;
  t.tm_mday = day;
  t.tm_hour = hour;
  t.tm_min = min;
  t.tm_sec = sec;
  return (int64_t)timegm(&t);
}

/* tcm:33-74, reading one line including its '\n'. */
static long read_line(char **buf, size_t *cap, FILE *fp) {
  if (*buf == NULL || *cap == 0) {
    *cap = 128;
    *buf = (char *)malloc(*cap);
    if (!*buf) return -1;
  }
  size_t i = 0;
  int c = 0;
  while ((c = fgetc(fp)) != EOF) {
    if (i + 1 >= *cap) {
      char *nb = (char *)realloc(*buf, *cap * 2);
      if (!nb) return -1;
      *buf = nb;
      *cap *= 2;
    }
    (*buf)[i++] = (char)c;
    if (c == '\n') break;
  }
  if (i == 0 && c == EOF) return -1;
  (*buf)[i] = '\0';
  return (long)i;
}

/* tcm:122-143 */
static int count_valid_edges(const char *path) {
  FILE *fp = fopen(path, "r");
  if (!fp) return -1;
  char *line = (char *)malloc(OR_MAX_LINE_LEN);
  int count = 0;
  while (fgets(line, OR_MAX_LINE_LEN, fp)) {
    if (line[0] == '#') continue;
    int d1, d2;
    char s[64], e[64];
    if (sscanf(line, " %d, %d, %63[^,], %63[^;];", &d1, &d2, s, e) == 4) count++;
  }
  free(line);
  fclose(fp);
  return count;
}

static or_graph *graph_alloc(int n_nodes, int cap_edges) {
  or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
  if (!g) return NULL;
  g->n_nodes = n_nodes;
  g->iv = (or_interval *)calloc((size_t)(cap_edges > 0 ? cap_edges : 1),
                                sizeof(or_interval));
  g->index_to_id = (int64_t *)calloc((size_t)(cap_edges > 0 ? 2 * cap_edges : 1),
                                     sizeof(int64_t));
  if (!g->iv || !g->index_to_id) {
    free(g->iv);
    free(g->index_to_id);
    free(g);
    return NULL;
  }
  return g;
}

void or_graph_free(or_graph *g) {
  if (!g) return;
  free(g->iv);
  free(g->sub);
  free(g->index_to_id);
  free(g);
}

int or_graph_n_nodes(const or_graph *g) { return g->n_nodes; }
int or_graph_n_edges(const or_graph *g) { return g->n_edges; }
int or_graph_n_top(const or_graph *g) { return g->n_top; }
int or_graph_nlists(const or_graph *g) { return g->nlists; }
const or_interval *or_graph_intervals(const or_graph *g) { return g->iv; }

/* Appends one valid contact (tcm:294-300). */
static int add_contact(or_graph *g, idmap *m, int src_id, int tgt_id,
                       int64_t start, int64_t end) {
  size_t before = m->count;
  int si = idmap_get_or_create(m, src_id);
  if (si < 0) return OR_E_NOMEM;
  if (m->count != before) g->index_to_id[si] = src_id;
  before = m->count;
  int ti = idmap_get_or_create(m, tgt_id);
  if (ti < 0) return OR_E_NOMEM;
  if (m->count != before) g->index_to_id[ti] = tgt_id;
  if ((int)m->count > g->n_nodes) return OR_E_NODE_RANGE;
  or_interval *iv = &g->iv[g->n_edges++];
  iv->start = start;
  iv->end = end;
  iv->target_id = or_pack_pair(si, ti);
  iv->sublist = -1;
  return OR_OK;
}

/* tcm:148-321 */
int or_parse_file(const char *path, or_graph **out) {
  *out = NULL;
  FILE *fp = fopen(path, "r");
  if (!fp) return OR_E_IO;
  char *header = NULL;
  size_t hcap = 0;
  long hlen = read_line(&header, &hcap, fp);
  fclose(fp);
  if (hlen <= 0) {
    free(header);
    return OR_E_HEADER;
  }
  if (strncmp(header, "NODE_LIST", 9) != 0 ||
      (header[9] != ' ' && header[9] != '\t')) {
    free(header);
    return OR_E_HEADER;
  }
  /* tcm:180-232: strtok ids, strtol-validated, duplicates fatal. */
  int n_ids = 0, id_cap = 1024;
  int *ids = (int *)malloc((size_t)id_cap * sizeof(int));
  char *save = NULL;
  char *tok = strtok_r(header + 9, " \t\r\n", &save);
  int rc = OR_OK;
  idmap dup;
  idmap_init(&dup, 1024);
  while (tok) {
    errno = 0;
    char *endp = NULL;
    long id = strtol(tok, &endp, 10);
    if (errno != 0 || *endp != '\0' || id < 0) {
      rc = OR_E_NODE_ID;
      break;
    }
    size_t before = dup.count;
    idmap_get_or_create(&dup, (int64_t)(int)id);
    if (dup.count == before) {
      rc = OR_E_DUP_ID;
      break;
    }
    if (n_ids >= id_cap) {
      id_cap *= 2;
      ids = (int *)realloc(ids, (size_t)id_cap * sizeof(int));
    }
    ids[n_ids++] = (int)id;
    tok = strtok_r(NULL, " \t\r\n", &save);
  }
  idmap_free(&dup);
  free(header);
  free(ids);
  if (rc != OR_OK) return rc;
  if (n_ids <= 0) return OR_E_NO_NODES;

  int edge_count = count_valid_edges(path);
  if (edge_count <= 0) return OR_E_NO_EDGES;
  fp = fopen(path, "r");
  if (!fp) return OR_E_IO;
  or_graph *g = graph_alloc(n_ids, edge_count);
  if (!g) {
    fclose(fp);
    return OR_E_NOMEM;
  }
  idmap m;
  idmap_init(&m, (size_t)n_ids);
  char *line = (char *)malloc(OR_MAX_LINE_LEN);
  while (fgets(line, OR_MAX_LINE_LEN, fp)) {
    if (line[0] == '#') continue;
    int src_id, tgt_id;
    char s[64], e[64];
    if (sscanf(line, " %d , %d , %63[^,] , %63[^;\r\n] ;", &src_id, &tgt_id, s,
               e) != 4)
      continue; /* tcm:280-283, warning only */
    int64_t start = or_parse_iso8601(s);
    int64_t end = or_parse_iso8601(e);
    if (start == -1 || end == -1) continue; /* tcm:285-288 */
    if (g->n_edges >= edge_count) { /* tcm:308-312 would exit() */
      rc = OR_E_COUNT;
      break;
    }
    rc = add_contact(g, &m, src_id, tgt_id, start, end);
    if (rc != OR_OK) break;
  }
  free(line);
  fclose(fp);
  g->n_indexed = (int)m.count;
  idmap_free(&m);
  if (rc == OR_OK && g->n_edges == 0) rc = OR_E_NO_EDGES; /* cu:735-738 */
  if (rc != OR_OK) {
    or_graph_free(g);
    return rc;
  }
  *out = g;
  return OR_OK;
}

int or_graph_from_contacts(int n_list, const int64_t *node_list, int n_contacts,
                           const int32_t *src_id, const int32_t *tgt_id,
                           const int64_t *start, const int64_t *end,
                           or_graph **out) {
  *out = NULL;
  if (n_list <= 0) return OR_E_NO_NODES;
  idmap dup;
  idmap_init(&dup, (size_t)n_list);
  for (int i = 0; i < n_list; i++) {
    if (node_list && node_list[i] < 0) {
      idmap_free(&dup);
      return OR_E_NODE_ID;
    }
    size_t before = dup.count;
    idmap_get_or_create(&dup, node_list ? node_list[i] : i);
    if (dup.count == before) {
      idmap_free(&dup);
      return OR_E_DUP_ID;
    }
  }
  idmap_free(&dup);
  if (n_contacts <= 0) return OR_E_NO_EDGES;
  or_graph *g = graph_alloc(n_list, n_contacts);
  if (!g) return OR_E_NOMEM;
  idmap m;
  idmap_init(&m, (size_t)n_list);
  int rc = OR_OK;
  for (int i = 0; i < n_contacts && rc == OR_OK; i++) {
    if (start[i] == -1 || end[i] == -1) continue; /* tcm:287 */
    rc = add_contact(g, &m, src_id[i], tgt_id[i], start[i], end[i]);
  }
  g->n_indexed = (int)m.count;
  idmap_free(&m);
  if (rc == OR_OK && g->n_edges == 0) rc = OR_E_NO_EDGES;
  if (rc != OR_OK) {
    or_graph_free(g);
    return rc;
  }
  *out = g;
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* NCLS: restates build_nested_list (idb:227-326) including the local       */
/* "nlists++ on every closed top-level chain" change (idb:258-261).         */
/* ------------------------------------------------------------------------ */

/* idb:30-46 */
static int start_cmp(const void *va, const void *vb) {
  const or_interval *a = (const or_interval *)va, *b = (const or_interval *)vb;
  if (a->start < b->start) return -1;
  if (a->start > b->start) return 1;
  if (a->end > b->end) return -1;
  if (a->end < b->end) return 1;
  return 0;
}

/* idb:69-83 as used on the temporary array: key (sublist, start). */
static int sub_cmp(const void *va, const void *vb) {
  const or_interval *a = (const or_interval *)va, *b = (const or_interval *)vb;
  if (a->sublist < b->sublist) return -1;
  if (a->sublist > b->sublist) return 1;
  if (a->start < b->start) return -1;
  if (a->start > b->start) return 1;
  return 0;
}

static int contains_not(const or_interval *x, const or_interval *p) {
  /* "i NOT CONTAINED" test of idb:247-249 */
  return x->end > p->end || (x->end == p->end && x->start == p->start);
}

/* idb:412-424 (ALL_POSITIVE_ORIENTATION is defined by intervaldb.h:195):
   an interval whose start sign differs from ori_sign is mirrored; the saved
   start goes through an int. */
static void reorient(int n, or_interval *im, int ori_sign) {
  for (int i = 0; i < n; i++) {
    if ((im[i].start >= 0 ? 1 : -1) != ori_sign) {
      int tmp = (int)im[i].start;
      im[i].start = -im[i].end;
      im[i].end = -tmp;
    }
  }
}

int or_build_ncls(or_graph *g) {
  int n = g->n_edges;
  or_interval *im = g->iv;
  if (n <= 0) return OR_E_NO_EDGES;
  for (int i = 0; i < n; i++) im[i].sublist = -1;
  reorient(n, im, 1); /* idb:233-235 */
  qsort(im, (size_t)n, sizeof(or_interval), start_cmp); /* idb:240 */
  int nsub = 0, nlists = 0;
  int i = 0;
  while (i < n) { /* idb:243-262 */
    int parent = i;
    i = parent + 1;
    while (i < n && parent >= 0) {
      if (contains_not(&im[i], &im[parent])) {
        parent = (int)im[parent].sublist;
      } else {
        im[i].sublist = parent;
        nsub++;
        parent = i;
        i++;
      }
    }
    if (parent < 0) nlists++;
  }
  or_sublist *sub = NULL;
  if (nsub > 0) { /* idb:265-314 */
    or_interval *tmp = (or_interval *)calloc((size_t)nsub, sizeof(or_interval));
    if (!tmp) return OR_E_NOMEM;
    int j = 0;
    for (i = 0; i < n; i++) {
      int parent = (int)im[i].sublist;
      if (parent >= 0) {
        tmp[j].start = i;
        tmp[j].sublist = parent;
        j++;
        if (im[parent].sublist < 0) im[parent].sublist = nlists++;
      }
      im[i].sublist = -1;
    }
    qsort(tmp, (size_t)nsub, sizeof(or_interval), sub_cmp);
    sub = (or_sublist *)calloc((size_t)nlists, sizeof(or_sublist));
    if (!sub) {
      free(tmp);
      return OR_E_NOMEM;
    }
    for (i = 0; i < nsub; i++) {
      int src = (int)tmp[i].start;
      int parent = (int)tmp[i].sublist;
      tmp[i] = im[src];
      int k = (int)im[parent].sublist;
      if (sub[k].len == 0) sub[k].start = i;
      sub[k].len++;
      im[src].start = im[src].end = -1;
    }
    j = 0;
    for (i = 0; i < n; i++) {
      if (im[i].start != -1 || im[i].end != -1) {
        if (j < i) im[j] = im[i];
        j++;
      }
    }
    memcpy(im + j, tmp, (size_t)nsub * sizeof(or_interval));
    for (i = 0; i < nlists; i++) sub[i].start += j;
    free(tmp);
    g->n_top = j;
  } else {
    sub = (or_sublist *)calloc(1, sizeof(or_sublist));
    if (!sub) return OR_E_NOMEM;
    g->n_top = n;
  }
  free(g->sub);
  g->sub = sub;
  g->nlists = nlists;
  g->ncls_built = 1;
  return OR_OK;
}

/* idb:344-360 */
static int64_t overlap_start(int64_t s, int64_t e, const or_interval *im, int n) {
  int64_t l = 0, r = n - 1;
  while (l < r) {
    int64_t mid = (l + r) / 2;
    if (im[mid].end <= s)
      l = mid + 1;
    else
      r = mid;
  }
  if (l < n && im[l].start < e && s < im[l].end) return l;
  return -1;
}

/* idb:426-485: depth-first over the nested lists, stopping once the hit
   buffer holds nbuf entries (the test is made after each store). */
int or_find_intervals(const or_graph *g, int64_t ws, int64_t we, int nbuf,
                      or_interval *buf, int *n_out) {
  const or_interval *im = g->iv;
  int depth_cap = 64, depth = 0;
  int64_t *st_i = (int64_t *)malloc(sizeof(int64_t) * (size_t)depth_cap);
  int64_t *st_n = (int64_t *)malloc(sizeof(int64_t) * (size_t)depth_cap);
  if (!st_i || !st_n) {
    free(st_i);
    free(st_n);
    return OR_E_NOMEM;
  }
  int64_t ibuf = 0;
  int ori_sign = 1;
  if (ws < 0) { /* idb:437-444: query mirrored into positive orientation */
    int64_t j = ws;
    ws = -we;
    we = -j;
    ori_sign = -1;
  }
  st_i[0] = overlap_start(ws, we, im, g->n_top);
  st_n[0] = g->n_top;
  for (;;) {
    int full = 0;
    while (st_i[depth] >= 0 && st_i[depth] < st_n[depth] &&
           im[st_i[depth]].start < we && ws < im[st_i[depth]].end) {
      const or_interval *hit = &im[st_i[depth]];
      buf[ibuf++] = *hit;
      int64_t k = hit->sublist;
      st_i[depth]++;
      if (k >= 0) {
        int64_t j = -1;
        const or_sublist *sh = &g->sub[k];
        int64_t jj = overlap_start(ws, we, im + sh->start, sh->len);
        if (jj >= 0) j = jj + sh->start;
        if (j >= 0) {
          if (depth + 1 >= depth_cap) {
            depth_cap *= 2;
            st_i = (int64_t *)realloc(st_i, sizeof(int64_t) * (size_t)depth_cap);
            st_n = (int64_t *)realloc(st_n, sizeof(int64_t) * (size_t)depth_cap);
          }
          depth++;
          st_i[depth] = j;
          st_n[depth] = sh->start + sh->len;
        }
      }
      if (ibuf >= nbuf) {
        full = 1;
        break;
      }
    }
    if (full || depth == 0) break;
    depth--;
  }
  free(st_i);
  free(st_n);
  reorient((int)ibuf, buf, ori_sign); /* idb:475-478 */
  *n_out = (int)ibuf;
  return OR_OK;
}

/* cu:973-989: buffer of max(nlists, n) entries, nbuf = nlists. */
int or_step_hits(const or_graph *g, int64_t ws, int64_t we, or_interval *buf,
                 int bufcap, int *n_out) {
  int need = g->nlists > 0 ? g->nlists : g->n_top;
  if (bufcap < need) return OR_E_PARAM;
  return or_find_intervals(g, ws, we, g->nlists, buf, n_out);
}

/* ------------------------------------------------------------------------ */
/* window arithmetic: cu:860-959                                              */
/* ------------------------------------------------------------------------ */
static int window_of(const or_params *p, int iter, int64_t *offset_out,
                     int64_t *ws_out) {
  double raw = (double)iter * (double)p->step_size;
  if (raw > (double)INT64_MAX || raw < (double)INT64_MIN) return OR_E_OVERFLOW;
  int64_t offset = (int64_t)llround(raw);
  if ((offset > 0 && p->start_time > INT64_MAX - offset) ||
      (offset < 0 && p->start_time < INT64_MIN - offset))
    return OR_E_OVERFLOW;
  int64_t D = p->static_network_duration;
  int64_t net = (offset / D) * D;
  if (offset < 0 && (offset % D)) net -= D;
  if ((net > 0 && p->start_time > INT64_MAX - net) ||
      (net < 0 && p->start_time < INT64_MIN - net))
    return OR_E_OVERFLOW;
  *offset_out = offset;
  *ws_out = p->start_time + net;
  return OR_OK;
}

int or_window(const or_params *p, or_window_cache *cache, int iter,
              int64_t *current_time, int64_t *ws, int64_t *we, float *scale) {
  if (p->static_network_duration <= 0 || !(p->step_size > 0.0f)) return OR_E_PARAM;
  int64_t offset, start_ts;
  int rc = window_of(p, iter, &offset, &start_ts);
  if (rc) return rc;
  if (start_ts > INT64_MAX - p->static_network_duration) return OR_E_OVERFLOW;
  if (!cache->valid || start_ts != cache->cached_start) {
    int steps = 0;
    for (int f = iter; f < p->iterations; ++f) {
      int64_t fo, fs;
      rc = window_of(p, f, &fo, &fs);
      if (rc) return rc;
      if (fs != start_ts) break;
      steps++;
    }
    if (steps <= 0) return OR_E_PARAM;
    cache->cached_start = start_ts;
    cache->cached_steps = steps;
    cache->valid = 1;
  }
  *current_time = p->start_time + offset;
  *ws = start_ts;
  *we = start_ts + p->static_network_duration;
  *scale = 1.0f / (float)cache->cached_steps;
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* CSR build: cu:238-350 (qsort by (row, col), merge by fp32 +=)             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int row;
  int col;
  float weight;
} edge_item;

static int edge_cmp(const void *va, const void *vb) { /* cu:216-229 */
  const edge_item *a = (const edge_item *)va, *b = (const edge_item *)vb;
  if (a->row < b->row) return -1;
  if (a->row > b->row) return 1;
  if (a->col < b->col) return -1;
  if (a->col > b->col) return 1;
  return 0;
}

int or_build_csr(const or_interval *hits, int n_hits, int n_nodes, int64_t ws,
                 int64_t we, float weight_scale, int32_t *rowptr,
                 int32_t *colind, float *val, int64_t *nnz) {
  memset(rowptr, 0, sizeof(int32_t) * (size_t)(n_nodes + 1));
  if (n_hits <= 0) {
    *nnz = 0;
    return OR_OK;
  }
  edge_item *el = (edge_item *)malloc(sizeof(edge_item) * 2 * (size_t)n_hits);
  if (!el) return OR_E_NOMEM;
  int ec = 0;
  for (int i = 0; i < n_hits; i++) {
    int src, tgt;
    unpack_pair(hits[i].target_id, &src, &tgt);
    int64_t os = hits[i].start > ws ? hits[i].start : ws;
    int64_t oe = hits[i].end < we ? hits[i].end : we;
    if (oe > os) {
      float w = (float)(oe - os) * weight_scale;
      el[ec].row = src;
      el[ec].col = tgt;
      el[ec].weight = w;
      ec++;
      el[ec].row = tgt;
      el[ec].col = src;
      el[ec].weight = w;
      ec++;
    }
  }
  qsort(el, (size_t)ec, sizeof(edge_item), edge_cmp);
  int agg = 0;
  for (int i = 1; i < ec; i++) {
    if (el[i].row == el[agg].row && el[i].col == el[agg].col) {
      el[agg].weight += el[i].weight;
    } else {
      ++agg;
      el[agg] = el[i];
    }
  }
  ec = ec > 0 ? agg + 1 : 0;
  for (int i = 0; i < ec; i++) rowptr[el[i].row + 1] += 1;
  for (int i = 0; i < n_nodes; i++) rowptr[i + 1] += rowptr[i];
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
  memcpy(pos, rowptr, sizeof(int32_t) * (size_t)n_nodes);
  for (int i = 0; i < ec; i++) {
    int r = el[i].row;
    colind[pos[r]] = el[i].col;
    val[pos[r]] = el[i].weight;
    pos[r]++;
  }
  free(pos);
  free(el);
  *nnz = ec;
  return OR_OK;
}

int or_step_csr(const or_graph *g, const or_params *p, int iter, int32_t *rowptr,
                int32_t *colind, float *val, int64_t *nnz, int64_t *ws_out,
                int64_t *we_out, float *scale_out) {
  /* the divisor is fixed at the first iteration of each window (cu:904-957),
     so the cache is replayed from iteration 0 */
  or_window_cache cache = {0, 0, 1};
  int64_t t, ws, we;
  float scale;
  int rc = OR_OK;
  for (int it = 0; it <= iter && rc == OR_OK; it++)
    rc = or_window(p, &cache, it, &t, &ws, &we, &scale);
  if (rc) return rc;
  int cap = g->nlists > 0 ? g->nlists : g->n_top;
  or_interval *hits = (or_interval *)malloc(sizeof(or_interval) * (size_t)cap);
  if (!hits) return OR_E_NOMEM;
  int nh = 0;
  rc = or_step_hits(g, ws, we, hits, cap, &nh);
  if (rc == OR_OK)
    rc = or_build_csr(hits, nh, g->n_nodes, ws, we, scale, rowptr, colind, val, nnz);
  free(hits);
  if (ws_out) *ws_out = ws;
  if (we_out) *we_out = we;
  if (scale_out) *scale_out = scale;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key  */
/* increments 0x9E3779B9 / 0xBB67AE85.                                      */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

uint32_t or_philox_word(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                        uint32_t c3) {
  uint32_t ctr[4] = {c0, c1, c2, c3};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  return out[0];
}

/* Uniform for (slot, node, replica): slot 0 = initial draw, slot k+1 = draw
   of iteration k.  One Philox4x32-10 block serves four replicas: replica r
   takes output word (r >> 6) & 3 of counter {node, (r >> 8) * 64 + (r & 63),
   slot, stream} under key = seed; its top 24 bits scaled by 2^-24 give u in
   [0, 1).  An injected stream is indexed [slot][replica][node]. */
float or_uniform(const or_params *p, int slot, int node, int replica, int n_nodes) {
  if (p->uniforms)
    return p->uniforms[((size_t)slot * (size_t)p->M + (size_t)replica) * (size_t)n_nodes +
                       (size_t)node];
  const uint32_t r = (uint32_t)replica;
  uint32_t ctr[4] = {(uint32_t)node, ((r >> 8) << 6) | (r & 63u), (uint32_t)slot, p->stream};
  uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  return (float)(out[(r >> 6) & 3u] >> 8) * 0x1p-24f;
}

/* ------------------------------------------------------------------------ */
/* or_expf: exp(x) evaluated in double and rounded once to float.           */
/* Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, then the degree-13     */
/* Taylor polynomial of e^r in Horner form with fused multiply-adds,        */
/* scaled by 2^k.  Relative error before the final rounding is below 1e-15, */
/* so the float result is the correctly rounded e^x except when e^x lies    */
/* within ~1e-15 of a rounding midpoint.                                     */
/* ------------------------------------------------------------------------ */
float or_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -104.0f) return 0.0f;
  const double xd = (double)x;
  const double kd = rint(xd * 1.4426950408889634074);
  double r = fma(-kd, 6.93147180369123816490e-01, xd);
  r = fma(-kd, 1.90821492927058770002e-10, r);
  double p = 1.0 / 6227020800.0; /* 1/13! */
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return (float)ldexp(p, (int)kd);
}

/* ------------------------------------------------------------------------ */
/* CPU twin (cu:358-500)                                                    */
/* ------------------------------------------------------------------------ */
static int32_t upper_range(const or_params *p) {
  return p->exposed_duration + p->infectious_duration + p->resistant_duration;
}
static int32_t medium_range(const or_params *p) {
  return p->infectious_duration + p->resistant_duration;
}
static int32_t lower_range(const or_params *p) { return p->resistant_duration; }

/* cu:362-374 */
void or_init_countdown(const or_params *p, int n_nodes, int32_t *c) {
  const int32_t med = medium_range(p);
  for (int r = 0; r < p->M; r++)
    for (int i = 0; i < n_nodes; i++) {
      size_t idx = (size_t)i + (size_t)r * (size_t)n_nodes;
      float u = or_uniform(p, 0, i, r, n_nodes);
      c[idx] = (u < p->initial_infected) ? (int32_t)((float)med + p->step_size) : 0;
    }
}

/* cu:381-394 */
void or_infectious_vector(const or_params *p, int n_nodes, const int32_t *c,
                          float *inf) {
  const int32_t lo = lower_range(p), med = medium_range(p);
  size_t total = (size_t)n_nodes * (size_t)p->M;
  for (size_t idx = 0; idx < total; idx++)
    inf[idx] = (c[idx] > lo && c[idx] <= med) ? 1.0f : 0.0f;
}

/* Eigen SparseDenseProduct.h:34-70 (RowMajor lhs, column-by-column): for each
   column, each row accumulates tmp += v * B(j, col) in ascending j from 0,
   then res = 0 + 1 * tmp.  Rows run under OpenMP when threads > 1 and
   nnz > 20000 (SparseDenseProduct.h:49). */
void or_spmm(const int32_t *rowptr, const int32_t *colind, const float *val,
             int64_t nnz, int n, const float *B, float *C, int m, int threads) {
  for (int col = 0; col < m; col++) {
    const float *b = B + (size_t)col * (size_t)n;
    float *cc = C + (size_t)col * (size_t)n;
#ifdef _OPENMP
    if (threads > 1 && nnz > 20000) {
      int chunk = (n + threads * 4 - 1) / (threads * 4);
#pragma omp parallel for schedule(dynamic, chunk) num_threads(threads)
      for (int i = 0; i < n; i++) {
        float tmp = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) tmp += val[j] * b[colind[j]];
        cc[i] = 0.0f + 1.0f * tmp;
      }
      continue;
    }
#else
    (void)threads;
    (void)nnz;
#endif
    for (int i = 0; i < n; i++) {
      float tmp = 0.0f;
      for (int j = rowptr[i]; j < rowptr[i + 1]; j++) tmp += val[j] * b[colind[j]];
      cc[i] = 0.0f + 1.0f * tmp;
    }
  }
}

void or_spmm_f64(const int32_t *rowptr, const int32_t *colind, const float *val,
                 int n, const float *B, double *C, int m) {
  for (int col = 0; col < m; col++)
    for (int i = 0; i < n; i++) {
      double tmp = 0.0;
      for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
        tmp += (double)val[j] * (double)B[(size_t)col * n + colind[j]];
      C[(size_t)col * n + i] = tmp;
    }
}

/* Seconds of contact per unit of infect_prob: the reference normalises the
   exposure to hours (cu:127-128, 449-450), whatever the step size.  A
   compile-time macro only so that tests/test_oracle_probability.py can build
   a copy with the reference's commented-out E / step_size (cu:449) and show
   that its closed-form checks reject it. */
#ifndef OR_HOUR_DIVISOR
#define OR_HOUR_DIVISOR(p) 3600.0f
#endif

/* cu:439-458 */
void or_infection_probability(const or_params *p, int n_nodes, const int32_t *c,
                              float *E) {
  const float log_base = logf(1.0f - p->infect_prob);
  size_t total = (size_t)n_nodes * (size_t)p->M;
  for (size_t idx = 0; idx < total; idx++) {
    if (c[idx] == 0) {
      float e = E[idx] / OR_HOUR_DIVISOR(p);
      float x = log_base * e;
      E[idx] = 1.0f - (p->use_libm_exp ? expf(x) : or_expf(x));
    } else {
      E[idx] = 0.0f;
    }
  }
}

/* cu:464-483 */
void or_monte_carlo(const or_params *p, int slot, int n_nodes, const float *P,
                    int32_t *c) {
  const int32_t up = upper_range(p);
  for (int r = 0; r < p->M; r++)
    for (int i = 0; i < n_nodes; i++) {
      size_t idx = (size_t)i + (size_t)r * (size_t)n_nodes;
      if (c[idx] == 0) {
        float u = or_uniform(p, slot, i, r, n_nodes);
        if (P[idx] > u) c[idx] = (int32_t)((float)up + p->step_size);
      }
    }
}

/* cu:489-500 */
void or_update_countdown(const or_params *p, int n_nodes, int32_t *c) {
  size_t total = (size_t)n_nodes * (size_t)p->M;
  for (size_t idx = 0; idx < total; idx++) {
    if (c[idx] > 0) c[idx] = (int32_t)((float)c[idx] - p->step_size);
    if (c[idx] < 0) c[idx] = 0;
  }
}

/* cu:530-550 */
int or_state_class(const or_params *p, int32_t v) {
  const int32_t up = upper_range(p);
  const int32_t Ed = p->exposed_duration, Id = p->infectious_duration,
                Rd = p->resistant_duration;
  int32_t elapsed = (v <= 0 || v > up) ? 0 : (up - v);
  if (v == 0) return 0;
  if (Ed > 0 && elapsed < Ed) return 1;
  if (Id > 0 && elapsed < Ed + Id) return 2;
  if (Rd > 0 && elapsed < Ed + Id + Rd) return 3;
  return 0;
}

void or_classify_counts(const or_params *p, int n_nodes, const int32_t *c,
                        int32_t *counts4) {
  for (int r = 0; r < p->M; r++) {
    int32_t k[4] = {0, 0, 0, 0};
    for (int i = 0; i < n_nodes; i++) k[or_state_class(p, c[(size_t)i + (size_t)r * n_nodes])]++;
    counts4[r * 4 + 0] = k[1];
    counts4[r * 4 + 1] = k[2];
    counts4[r * 4 + 2] = k[3];
    counts4[r * 4 + 3] = k[0];
  }
}

static const char *state_name(int cls) {
  static const char *names[4] = {"susceptible", "incubating", "infectious", "resistant"};
  return names[cls];
}

/* cu:505-568 */
static void write_state(const or_params *p, int n_nodes, const int32_t *c,
                        int64_t t, FILE *summary, FILE *node) {
  for (int r = 0; r < p->M; r++) {
    int32_t k[4] = {0, 0, 0, 0};
    for (int i = 0; i < n_nodes; i++) {
      int cls = or_state_class(p, c[(size_t)i + (size_t)r * n_nodes]);
      k[cls]++;
      if (node) fprintf(node, "%lld,%d,%d,%s\n", (long long)t, r, i, state_name(cls));
    }
    if (summary)
      fprintf(summary, "%lld,%d,%d,%d,%d,%d\n", (long long)t, r, k[1], k[2], k[3], k[0]);
  }
  if (summary) fflush(summary);
  if (node) fflush(node);
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* cu:727-1150 */
int or_simulate(const or_graph *g, const or_params *p, int32_t *counts,
                int32_t *final_state, FILE *summary, FILE *node,
                double *loop_seconds) {
  if (!g->ncls_built) return OR_E_PARAM;
  if (p->M <= 0 || p->iterations < 0) return OR_E_PARAM;
  if (p->static_network_duration <= 0 || !(p->step_size > 0.0f)) return OR_E_PARAM;
  const int N = g->n_nodes, M = p->M;
  const size_t total = (size_t)N * (size_t)M;
  int32_t *c = (int32_t *)malloc(total * sizeof(int32_t));
  float *E = (float *)malloc(total * sizeof(float));
  float *I = (float *)malloc(total * sizeof(float));
  int cap = g->nlists > 0 ? g->nlists : g->n_top;
  or_interval *hits = (or_interval *)malloc(sizeof(or_interval) * (size_t)cap);
  int32_t *rowptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N + 1));
  int32_t *colind = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)cap);
  float *val = (float *)malloc(sizeof(float) * 2 * (size_t)cap);
  int rc = OR_OK;
  if (!c || !E || !I || !hits || !rowptr || !colind || !val) {
    rc = OR_E_NOMEM;
    goto done;
  }
  or_init_countdown(p, N, c);
  if (counts) or_classify_counts(p, N, c, counts);
  if (summary || node) write_state(p, N, c, p->start_time, summary, node);
  double t0 = now_seconds();
  or_window_cache cache = {0, 0, 1};
  for (int iter = 0; iter < p->iterations; iter++) {
    int64_t t, ws, we;
    float scale;
    rc = or_window(p, &cache, iter, &t, &ws, &we, &scale);
    if (rc) goto done;
    int nh = 0;
    rc = or_step_hits(g, ws, we, hits, cap, &nh);
    if (rc) goto done;
    int64_t nnz = 0;
    rc = or_build_csr(hits, nh, N, ws, we, scale, rowptr, colind, val, &nnz);
    if (rc) goto done;
    or_infectious_vector(p, N, c, I);
    or_spmm(rowptr, colind, val, nnz, N, I, E, M, p->threads);
    or_infection_probability(p, N, c, E);
    or_monte_carlo(p, iter + 1, N, E, c);
    or_update_countdown(p, N, c);
    if (counts) or_classify_counts(p, N, c, counts + (size_t)(iter + 1) * M * 4);
    if (summary || node) write_state(p, N, c, t, summary, node);
  }
  if (loop_seconds) *loop_seconds = now_seconds() - t0;
  if (final_state) memcpy(final_state, c, total * sizeof(int32_t));
done:
  free(c);
  free(E);
  free(I);
  free(hits);
  free(rowptr);
  free(colind);
  free(val);
  return rc;
}

int or_simulate_csv(const or_graph *g, const or_params *p,
                    const char *summary_path, const char *node_path,
                    int32_t *counts, double *loop_seconds) {
  FILE *s = NULL, *n = NULL;
  if (summary_path) {
    remove(summary_path);
    s = fopen(summary_path, "w");
    if (!s) return OR_E_IO;
    fprintf(s, "time,simulation_id,exposed,infectious,resistant,susceptible\n");
  }
  if (node_path) {
    remove(node_path);
    n = fopen(node_path, "w");
    if (!n) {
      if (s) fclose(s);
      return OR_E_IO;
    }
    fprintf(n, "time,simulation_id,node_id,state\n");
  }
  int rc = or_simulate(g, p, counts, NULL, s, n, loop_seconds);
  if (s) fclose(s);
  if (n) fclose(n);
  return rc;
}
