/*
 * CPU oracle for the exact Cayley-graph BFS path.  TEST INFRASTRUCTURE ONLY:
 * tests/, the smoke check and the CPU-baseline leg of bench.py may load this
 * library as the checker; the HIP engine never links, calls or falls back to it.
 *
 * Each function restates the reference's CPU semantics (citations are paths in
 * the TryDotAtwo/MultiGPUBFS checkout):
 *
 *  orc_successor         multigpubfs/graph_definition.py:70-83 (permutation
 *                        gather child[i]=parent[g[i]]; matrix product with
 *                        wrapping int64 sum, then Python/rem_euclid modulus),
 *                        crates/mgbfs-core/src/graph_definition.rs:50-61.
 *  orc_exact_layers      multigpubfs/graph_definition.py:109-117 (level-
 *                        synchronous BFS over ordered sets; each layer sorted
 *                        lexicographically, capacity check ORACLE_CAPACITY).
 *  orc_pack/unpack,      scripts/state_layout.py:9-16 (4-bit symbols) and
 *  orc_packed_apply      gpu/cayley_bfs.cu:50-56 (state = sum perm[i] << 4i).
 *  orc_lrx_transitions   scripts/cpu_small_bfs.py:21-27 (packed L, R, X).
 *  orc_rank/unrank       rust/src/cayley.rs:504-516 (lexicographic Lehmer rank,
 *                        position 0 most significant) and
 *                        crates/mgbfs-runtime/src/permutation_bitmap.rs:18-64
 *                        (multiset rank/unrank; distinct symbols reduce to it).
 *  orc_packed_bfs        scripts/cpu_small_bfs.py:30-80 (packed-word frontier
 *                        loop) with the exact Lehmer-bitmap visited set of
 *                        gpu/cayley_bfs.cu:58-124, multi-threaded.
 *
 * Pinning (tests/test_oracle.py): reference golden layer tables for LRX S10
 * and S13, rank round trips at the reference's KAT ranks, exhaustive
 * lexicographic order, Mahonian layer sizes for adjacent transpositions, the
 * reference's action fixtures, and |SL(2,Z_p)| = p(p^2-1).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ARGUMENT 1
#define ORC_CAPACITY 2
#define ORC_MEMORY 3

/* ---------------------------------------------------------------- actions */

static int64_t wrap_mul_add(int64_t acc, int64_t a, int64_t b) {
  return (int64_t)((uint64_t)acc + (uint64_t)a * (uint64_t)b);
}

static int64_t python_mod(int64_t x, int64_t m) {
  int64_t r = x % m;
  if (r < 0) r += m;
  return r;
}

/* kind 0: permutation (elements = degree), kind 1: matrix rows x cols. */
int orc_successor(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols,
                  const uint32_t* perm, const int64_t* matrix, uint32_t modulus,
                  const int64_t* state, int64_t* out) {
  if (kind == 0) {
    for (uint32_t i = 0; i < elements; ++i) {
      if (perm[i] >= elements) return ORC_ARGUMENT;
      out[i] = state[perm[i]];
    }
    return ORC_OK;
  }
  if (kind != 1 || rows * cols != elements) return ORC_ARGUMENT;
  for (uint32_t i = 0; i < rows; ++i)
    for (uint32_t j = 0; j < cols; ++j) {
      int64_t total = 0;
      for (uint32_t k = 0; k < rows; ++k)
        total = wrap_mul_add(total, matrix[i * rows + k], state[k * cols + j]);
      out[i * cols + j] = modulus ? python_mod(total, (int64_t)modulus) : total;
    }
  return ORC_OK;
}

/* ------------------------------------------------- ordered-set exact layers */

static uint32_t g_cmp_width;
static int cmp_rows(const void* a, const void* b) {
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  for (uint32_t i = 0; i < g_cmp_width; ++i) {
    if (x[i] < y[i]) return -1;
    if (x[i] > y[i]) return 1;
  }
  return 0;
}
static int rows_cmp(const int64_t* x, const int64_t* y, uint32_t w) {
  for (uint32_t i = 0; i < w; ++i) {
    if (x[i] < y[i]) return -1;
    if (x[i] > y[i]) return 1;
  }
  return 0;
}
static int contains_sorted(const int64_t* rows, uint64_t count, uint32_t w, const int64_t* key) {
  uint64_t lo = 0, hi = count;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    int c = rows_cmp(rows + mid * w, key, w);
    if (c == 0) return 1;
    if (c < 0) lo = mid + 1; else hi = mid;
  }
  return 0;
}

/*
 * Exact layered BFS over ordered sets.  `states_out` receives every layer's
 * states concatenated (layer d sorted lexicographically), capacity
 * `max_states` rows of `elements` int64 values; `layer_sizes` receives up to
 * `max_layers` sizes.  Returns ORC_CAPACITY when |seen| would exceed
 * max_states (graph_definition.py:115) or when max_layers is too small.
 */
int orc_exact_layers(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols,
                     uint32_t generators, const uint32_t* perms, const int64_t* matrices,
                     const uint32_t* moduli, const int64_t* start, uint64_t max_states,
                     int64_t* states_out, uint64_t* layer_sizes, uint32_t max_layers,
                     uint32_t* layer_count) {
  if (!elements || !generators || !max_states || !max_layers) return ORC_ARGUMENT;
  const uint32_t w = elements;
  int64_t* seen = (int64_t*)malloc(sizeof(int64_t) * w * (max_states + 1));
  int64_t* merged = (int64_t*)malloc(sizeof(int64_t) * w * (max_states + 1));
  int64_t* future = NULL;
  uint64_t seen_n = 1, frontier_n = 1, written = 0;
  uint32_t depth = 0;
  int status = ORC_OK;
  if (!seen || !merged) { free(seen); free(merged); return ORC_MEMORY; }
  memcpy(seen, start, sizeof(int64_t) * w);
  memcpy(states_out, start, sizeof(int64_t) * w);
  written = 1;
  layer_sizes[0] = 1;
  depth = 1;
  g_cmp_width = w;
  while (frontier_n) {
    const int64_t* frontier = states_out + (written - frontier_n) * w;
    uint64_t cand_n = frontier_n * generators;
    future = (int64_t*)malloc(sizeof(int64_t) * w * cand_n);
    if (!future) { status = ORC_MEMORY; break; }
    uint64_t c = 0;
    for (uint64_t p = 0; p < frontier_n; ++p)
      for (uint32_t g = 0; g < generators; ++g) {
        int64_t* out = future + c * w;
        int rc = orc_successor(kind, w, rows, cols, kind == 0 ? perms + (uint64_t)g * w : NULL,
                               kind == 1 ? matrices + (uint64_t)g * rows * rows : NULL,
                               kind == 1 ? moduli[g] : 0, frontier + p * w, out);
        if (rc) { status = rc; goto done; }
        if (!contains_sorted(seen, seen_n, w, out)) ++c;
      }
    qsort(future, c, sizeof(int64_t) * w, cmp_rows);
    uint64_t u = 0;
    for (uint64_t i = 0; i < c; ++i)
      if (u == 0 || rows_cmp(future + (u - 1) * w, future + i * w, w) != 0) {
        if (u != i) memcpy(future + u * w, future + i * w, sizeof(int64_t) * w);
        ++u;
      }
    if (u == 0) break;
    if (seen_n + u > max_states) { status = ORC_CAPACITY; goto done; }
    if (depth >= max_layers) { status = ORC_CAPACITY; goto done; }
    memcpy(states_out + written * w, future, sizeof(int64_t) * w * u);
    written += u;
    layer_sizes[depth++] = u;
    /* merge sorted `seen` with sorted `future` */
    uint64_t i = 0, j = 0, k = 0;
    while (i < seen_n || j < u) {
      if (j >= u || (i < seen_n && rows_cmp(seen + i * w, future + j * w, w) < 0))
        memcpy(merged + (k++) * w, seen + (i++) * w, sizeof(int64_t) * w);
      else
        memcpy(merged + (k++) * w, future + (j++) * w, sizeof(int64_t) * w);
    }
    int64_t* t = seen; seen = merged; merged = t;
    seen_n = k;
    frontier_n = u;
    free(future);
    future = NULL;
  }
done:
  free(future);
  free(seen);
  free(merged);
  *layer_count = depth;
  return status;
}

/* ------------------------------------------------------ packed nibble keys */

uint64_t orc_pack(const int64_t* state, uint32_t n) {
  uint64_t key = 0;
  for (uint32_t i = 0; i < n; ++i) key |= ((uint64_t)state[i] & 15u) << (4 * i);
  return key;
}

void orc_unpack(uint64_t key, uint32_t n, int64_t* state) {
  for (uint32_t i = 0; i < n; ++i) state[i] = (int64_t)((key >> (4 * i)) & 15u);
}

uint64_t orc_packed_apply(uint64_t key, const uint32_t* perm, uint32_t n) {
  uint64_t child = 0;
  for (uint32_t i = 0; i < n; ++i) child |= ((key >> (4 * perm[i])) & 15u) << (4 * i);
  return child;
}

/* Packed L, R, X words (cpu_small_bfs.py:21-27), `bits` per symbol. */
void orc_lrx_transitions(uint64_t word, uint32_t n, uint32_t bits, uint64_t out[3]) {
  uint64_t symbol = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
  uint64_t mask = (n * bits >= 64) ? ~0ull : ((1ull << (n * bits)) - 1);
  uint64_t difference = (word ^ (word >> bits)) & symbol;
  out[0] = (word >> bits) | ((word & symbol) << ((n - 1) * bits));
  out[1] = ((word << bits) & mask) | (word >> ((n - 1) * bits));
  out[2] = word ^ difference ^ (difference << bits);
}

/* ------------------------------------------------------------------ ranks */

static uint64_t factorial(uint32_t n) {
  uint64_t f = 1;
  for (uint32_t i = 2; i <= n; ++i) f *= i;
  return f;
}

/* Lexicographic rank of a permutation of 0..n-1 (n <= 20). */
uint64_t orc_rank(const int64_t* perm, uint32_t n) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t smaller = 0;
    for (uint32_t j = i + 1; j < n; ++j) smaller += perm[j] < perm[i];
    r = r * (n - i) + smaller;
  }
  return r;
}

int orc_unrank(uint64_t rank, uint32_t n, int64_t* perm) {
  if (n == 0 || n > 20 || rank >= factorial(n)) return ORC_ARGUMENT;
  int used[20] = {0};
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t f = factorial(n - 1 - i);
    uint64_t d = rank / f;
    rank %= f;
    for (uint32_t s = 0; s < n; ++s) {
      if (used[s]) continue;
      if (d == 0) { perm[i] = s; used[s] = 1; break; }
      --d;
    }
  }
  return ORC_OK;
}

/* Multiset lexicographic rank (permutation_bitmap.rs:33-46), 128-bit safe. */
int orc_multiset_rank(const int64_t* state, uint32_t n, uint64_t* rank_out) {
  int64_t symbols[64];
  uint32_t counts[64];
  uint32_t k = 0;
  if (n == 0 || n > 64) return ORC_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t j = 0;
    while (j < k && symbols[j] != state[i]) ++j;
    if (j == k) { symbols[k] = state[i]; counts[k++] = 0; }
    counts[j]++;
  }
  /* sort symbols ascending with counts */
  for (uint32_t a = 1; a < k; ++a)
    for (uint32_t b = a; b > 0 && symbols[b - 1] > symbols[b]; --b) {
      int64_t s = symbols[b]; symbols[b] = symbols[b - 1]; symbols[b - 1] = s;
      uint32_t c = counts[b]; counts[b] = counts[b - 1]; counts[b - 1] = c;
    }
  unsigned __int128 ways = 1, prev = 0;
  for (uint32_t s = 0; s < k; ++s) {
    for (uint32_t t = 1; t <= counts[s]; ++t) {
      ways = ways * (prev + t) / t;
      if (ways > (unsigned __int128)UINT64_MAX) return ORC_CAPACITY;
    }
    prev += counts[s];
  }
  unsigned __int128 rank = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t s = 0;
    while (symbols[s] != state[i]) ++s;
    unsigned __int128 less = 0;
    for (uint32_t t = 0; t < s; ++t) less += counts[t];
    unsigned __int128 left = n - i;
    rank += ways * less / left;
    ways = ways * counts[s] / left;
    counts[s]--;
  }
  *rank_out = (uint64_t)rank;
  return ORC_OK;
}

/* --------------------------------------------- packed BFS, Lehmer bitmap */

static uint64_t packed_rank(uint64_t key, uint32_t n) {
  uint32_t seen = 0;
  uint64_t r = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t s = (uint32_t)((key >> (4 * i)) & 15u);
    uint32_t d = s - (uint32_t)__builtin_popcount(seen & ((1u << s) - 1u));
    seen |= 1u << s;
    r = r * (n - i) + d;
  }
  return r;
}

typedef struct {
  const uint64_t* frontier;
  uint64_t begin, end;
  uint32_t n, generators;
  const uint32_t* perms;
  uint64_t* bitmap;
  uint64_t* out;
  uint64_t out_n;
  uint64_t out_cap;
  int lrx;
} orc_task;

static void* packed_worker(void* arg) {
  orc_task* t = (orc_task*)arg;
  uint64_t children[64];
  for (uint64_t p = t->begin; p < t->end; ++p) {
    uint64_t key = t->frontier[p];
    if (t->lrx) {
      orc_lrx_transitions(key, t->n, 4, children);
    } else {
      for (uint32_t g = 0; g < t->generators; ++g)
        children[g] = orc_packed_apply(key, t->perms + (uint64_t)g * t->n, t->n);
    }
    for (uint32_t g = 0; g < t->generators; ++g) {
      uint64_t r = packed_rank(children[g], t->n);
      uint64_t bit = 1ull << (r & 63);
      uint64_t old = __atomic_fetch_or(&t->bitmap[r >> 6], bit, __ATOMIC_RELAXED);
      if (old & bit) continue;
      if (t->out_n == t->out_cap) {
        uint64_t cap = t->out_cap ? 2 * t->out_cap : 1024;
        uint64_t* grown = (uint64_t*)realloc(t->out, cap * sizeof(uint64_t));
        if (!grown) return (void*)1;
        t->out = grown;
        t->out_cap = cap;
      }
      t->out[t->out_n++] = children[g];
    }
  }
  return NULL;
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

/*
 * Packed BFS over permutations of 0..n-1 (n <= 16) with an n!-bit Lehmer
 * bitmap.  generators == 0 selects the packed LRX words of cpu_small_bfs.py.
 * layer_sizes: up to max_layers entries.  If keys_out != NULL every layer's
 * keys are written (each layer sorted ascending) up to keys_cap keys.
 * stop_after_edges > 0 stops after the layer at which that many edges have
 * been expanded (bounded CPU-baseline samples); *edges_out reports them.
 */
int orc_packed_bfs(uint32_t n, uint32_t generators, const uint32_t* perms, uint64_t start_key,
                   uint32_t threads, uint64_t* layer_sizes, uint32_t max_layers,
                   uint32_t* layer_count, uint64_t* keys_out, uint64_t keys_cap,
                   uint64_t stop_after_edges, uint64_t* edges_out) {
  if (n < 2 || n > 16 || threads == 0 || threads > 256 || !max_layers) return ORC_ARGUMENT;
  int lrx = generators == 0;
  uint32_t gens = lrx ? 3 : generators;
  if (gens > 64) return ORC_ARGUMENT;
  uint64_t domain = factorial(n);
  uint64_t words = (domain + 63) / 64;
  uint64_t* bitmap = (uint64_t*)calloc(words, sizeof(uint64_t));
  uint64_t* frontier = (uint64_t*)malloc(sizeof(uint64_t));
  if (!bitmap || !frontier) { free(bitmap); free(frontier); return ORC_MEMORY; }
  uint64_t r0 = packed_rank(start_key, n);
  bitmap[r0 >> 6] |= 1ull << (r0 & 63);
  frontier[0] = start_key;
  uint64_t frontier_n = 1, written = 0, edges = 0;
  uint32_t depth = 0;
  int status = ORC_OK;
  orc_task* tasks = (orc_task*)calloc(threads, sizeof(orc_task));
  pthread_t* tids = (pthread_t*)calloc(threads, sizeof(pthread_t));
  if (!tasks || !tids) { status = ORC_MEMORY; goto out; }
  while (frontier_n) {
    if (depth >= max_layers) { status = ORC_CAPACITY; break; }
    layer_sizes[depth++] = frontier_n;
    if (keys_out) {
      if (written + frontier_n > keys_cap) { status = ORC_CAPACITY; break; }
      memcpy(keys_out + written, frontier, frontier_n * sizeof(uint64_t));
      qsort(keys_out + written, frontier_n, sizeof(uint64_t), cmp_u64);
      written += frontier_n;
    }
    edges += frontier_n * gens;
    if (stop_after_edges && edges >= stop_after_edges) break;
    uint32_t used = threads;
    if (frontier_n < 4096) used = 1;
    uint64_t chunk = (frontier_n + used - 1) / used;
    for (uint32_t t = 0; t < used; ++t) {
      orc_task* k = &tasks[t];
      k->frontier = frontier;
      k->begin = t * chunk < frontier_n ? t * chunk : frontier_n;
      k->end = (t + 1) * chunk < frontier_n ? (t + 1) * chunk : frontier_n;
      k->n = n; k->generators = gens; k->perms = perms; k->bitmap = bitmap;
      k->out_n = 0; k->lrx = lrx;
    }
    if (used == 1) {
      if (packed_worker(&tasks[0])) { status = ORC_MEMORY; break; }
    } else {
      for (uint32_t t = 0; t < used; ++t) pthread_create(&tids[t], NULL, packed_worker, &tasks[t]);
      void* rv = NULL;
      int failed = 0;
      for (uint32_t t = 0; t < used; ++t) { pthread_join(tids[t], &rv); failed |= rv != NULL; }
      if (failed) { status = ORC_MEMORY; break; }
    }
    uint64_t next_n = 0;
    for (uint32_t t = 0; t < used; ++t) next_n += tasks[t].out_n;
    free(frontier);
    frontier = (uint64_t*)malloc((next_n ? next_n : 1) * sizeof(uint64_t));
    if (!frontier) { status = ORC_MEMORY; break; }
    uint64_t at = 0;
    for (uint32_t t = 0; t < used; ++t) {
      memcpy(frontier + at, tasks[t].out, tasks[t].out_n * sizeof(uint64_t));
      at += tasks[t].out_n;
    }
    frontier_n = next_n;
  }
out:
  if (tasks)
    for (uint32_t t = 0; t < threads; ++t) free(tasks[t].out);
  free(tasks);
  free(tids);
  free(frontier);
  free(bitmap);
  *layer_count = depth;
  if (edges_out) *edges_out = edges;
  return status;
}
