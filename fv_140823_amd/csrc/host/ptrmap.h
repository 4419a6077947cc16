/*
 * ptrmap.h — leaf pointer -> list position, for code that walks the leaf list of any RUN.
 *
 * The reference's BRANCH has no field for a leaf's position in the list, and a RUN built
 * by the reference's host code must not be written to beyond its own fields, so the
 * flattener and the export look positions up in this open-addressing table instead
 * (linear probing, at most half full, keys are the BRANCH addresses).
 */
#ifndef FV_PTRMAP_H
#define FV_PTRMAP_H

#include <stdint.h>
#include <stdlib.h>

typedef struct fv_ptrmap {
  const void **key;
  int32_t *val;
  size_t mask;
} fv_ptrmap;

static inline size_t fv_ptrmap_slot(const fv_ptrmap *m, const void *k) {
  uint64_t h = (uint64_t)(uintptr_t)k;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  return (size_t)h & m->mask;
}

/* room for n keys; 0 or -1 (no memory) */
static inline int fv_ptrmap_init(fv_ptrmap *m, size_t n) {
  size_t cap = 16;
  while (cap < 2 * n) cap <<= 1;
  m->key = (const void **)calloc(cap, sizeof(void *));
  m->val = (int32_t *)malloc(cap * sizeof(int32_t));
  m->mask = cap - 1;
  if (!m->key || !m->val) {
    free(m->key);
    free(m->val);
    m->key = NULL;
    m->val = NULL;
    return -1;
  }
  return 0;
}

static inline void fv_ptrmap_free(fv_ptrmap *m) {
  free(m->key);
  free(m->val);
  m->key = NULL;
  m->val = NULL;
}

static inline void fv_ptrmap_put(fv_ptrmap *m, const void *k, int32_t v) {
  size_t s = fv_ptrmap_slot(m, k);
  while (m->key[s] && m->key[s] != k) s = (s + 1) & m->mask;
  m->key[s] = k;
  m->val[s] = v;
}

/* the position of k, or -1 if k is not in the table */
static inline int32_t fv_ptrmap_get(const fv_ptrmap *m, const void *k) {
  if (!k) return -1;
  size_t s = fv_ptrmap_slot(m, k);
  while (m->key[s]) {
    if (m->key[s] == k) return m->val[s];
    s = (s + 1) & m->mask;
  }
  return -1;
}

#endif /* FV_PTRMAP_H */
