// Deterministic-mode rank sorters (device, one lane; the host orders barriers with the same
// pdq_sort).
//
// rank_lefs (src/libmodle/cpu/simulation.cpp:402-487) orders LEF indices by unit position with
// cpp-sort 1.13's pdq_sort (first epoch) or split_sort (later epochs), then re-sorts runs of
// equal positions by binding epoch with a stable insertion sort. Both sorters are unstable, so
// indices whose position and binding epoch tie keep whatever order the algorithm leaves; the
// deterministic mode reproduces that order by running the same algorithms:
//   pdq_sort   pattern-defeating quicksort: insertion sort below 24 elements, median-of-3
//              (ninther above 128), partition_right / partition_left, partial insertion sort on
//              already-partitioned ranges, pattern-breaking swaps, heapsort after log2(n) bad
//              partitions;
//   split_sort keep a non-decreasing run in front (an element that breaks it also evicts the
//              run's top), pdq-sort the evicted elements, merge the two runs stably.
#pragma once
#ifndef MGPU_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace mgpu {

struct KeyLess {
  const uint64_t* key;
  __host__ __device__ bool operator()(uint32_t a, uint32_t b) const { return key[a] < key[b]; }
};

__host__ __device__ inline void swap_u32(uint32_t* a, uint32_t* b) {
  const uint32_t t = *a;
  *a = *b;
  *b = t;
}

template <class L>
__host__ __device__ void ins_sort(uint32_t* begin, uint32_t* end, L less) {
  if (begin == end) return;
  for (uint32_t* cur = begin + 1; cur != end; ++cur) {
    uint32_t* sift = cur;
    uint32_t* sift_1 = cur - 1;
    if (less(*sift, *sift_1)) {
      const uint32_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (sift != begin && less(tmp, *--sift_1));
      *sift = tmp;
    }
  }
}

template <class L>
__host__ __device__ void unguarded_ins_sort(uint32_t* begin, uint32_t* end, L less) {
  if (begin == end) return;
  for (uint32_t* cur = begin + 1; cur != end; ++cur) {
    uint32_t* sift = cur;
    uint32_t* sift_1 = cur - 1;
    if (less(*sift, *sift_1)) {
      const uint32_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (less(tmp, *--sift_1));
      *sift = tmp;
    }
  }
}

template <class L>
__host__ __device__ bool partial_ins_sort(uint32_t* begin, uint32_t* end, L less) {
  if (begin == end) return true;
  uint64_t moved = 0;
  for (uint32_t* cur = begin + 1; cur != end; ++cur) {
    uint32_t* sift = cur;
    uint32_t* sift_1 = cur - 1;
    if (less(*sift, *sift_1)) {
      const uint32_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (sift != begin && less(tmp, *--sift_1));
      *sift = tmp;
      moved += (uint64_t)(cur - sift);
    }
    if (moved > 8) return false;
  }
  return true;
}

template <class L>
__host__ __device__ inline void sort3(uint32_t* a, uint32_t* b, uint32_t* c, L less) {
  if (less(*b, *a)) swap_u32(a, b);
  if (less(*c, *b)) swap_u32(b, c);
  if (less(*b, *a)) swap_u32(a, b);
}

template <class L>
__host__ __device__ uint32_t* partition_right(uint32_t* begin, uint32_t* end, L less, bool& already) {
  const uint32_t pivot = *begin;
  uint32_t* first = begin;
  uint32_t* last = end;
  while (less(*++first, pivot)) {
  }
  if (first - 1 == begin) {
    while (first < last && !less(*--last, pivot)) {
    }
  } else {
    while (!less(*--last, pivot)) {
    }
  }
  already = first >= last;
  while (first < last) {
    swap_u32(first, last);
    while (less(*++first, pivot)) {
    }
    while (!less(*--last, pivot)) {
    }
  }
  uint32_t* pivot_pos = first - 1;
  *begin = *pivot_pos;
  *pivot_pos = pivot;
  return pivot_pos;
}

template <class L>
__host__ __device__ uint32_t* partition_left(uint32_t* begin, uint32_t* end, L less) {
  const uint32_t pivot = *begin;
  uint32_t* first = begin;
  uint32_t* last = end;
  while (less(pivot, *--last)) {
  }
  if (last + 1 == end) {
    while (first < last && !less(pivot, *++first)) {
    }
  } else {
    while (!less(pivot, *++first)) {
    }
  }
  while (first < last) {
    swap_u32(first, last);
    while (less(pivot, *--last)) {
    }
    while (!less(pivot, *++first)) {
    }
  }
  uint32_t* pivot_pos = last;
  *begin = *pivot_pos;
  *pivot_pos = pivot;
  return pivot_pos;
}

// make_heap + sort_heap with the hole-sifting of libstdc++'s __adjust_heap/__push_heap.
template <class L>
__host__ __device__ void adjust_heap(uint32_t* first, int64_t hole, int64_t len, uint32_t value, L less) {
  const int64_t top = hole;
  int64_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (less(first[child], first[child - 1])) child--;
    first[hole] = first[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    first[hole] = first[child - 1];
    hole = child - 1;
  }
  int64_t parent = (hole - 1) / 2;
  while (hole > top && less(first[parent], value)) {
    first[hole] = first[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  first[hole] = value;
}

template <class L>
__host__ __device__ void heap_sort(uint32_t* first, uint32_t* last, L less) {
  const int64_t len = last - first;
  if (len >= 2) {
    for (int64_t parent = (len - 2) / 2;; --parent) {
      adjust_heap(first, parent, len, first[parent], less);
      if (parent == 0) break;
    }
  }
  while (last - first > 1) {
    --last;
    const uint32_t value = *last;
    *last = *first;
    adjust_heap(first, 0, last - first, value, less);
  }
}

// Iterative pdqsort: the left partition is sorted by an explicit stack instead of recursion.
template <class L>
__host__ __device__ void pdq_sort(uint32_t* base, uint32_t n, L less) {
  if (n == 0) return;
  int lg = 0;
  for (uint32_t m = n; m >>= 1;) ++lg;
  struct Frame {
    uint32_t* begin;
    uint32_t* end;
    int bad;
    bool leftmost;
  };
  Frame stack[64];
  int sp = 0;
  stack[sp++] = {base, base + n, lg, true};
  while (sp > 0) {
    Frame f = stack[--sp];
    uint32_t* begin = f.begin;
    uint32_t* end = f.end;
    int bad = f.bad;
    bool leftmost = f.leftmost;
    for (;;) {
      const int64_t size = end - begin;
      if (size < 24) {
        if (leftmost) ins_sort(begin, end, less);
        else unguarded_ins_sort(begin, end, less);
        break;
      }
      const int64_t s2 = size / 2;
      if (size > 128) {
        sort3(begin, begin + s2, end - 1, less);
        sort3(begin + 1, begin + (s2 - 1), end - 2, less);
        sort3(begin + 2, begin + (s2 + 1), end - 3, less);
        sort3(begin + (s2 - 1), begin + s2, begin + (s2 + 1), less);
        swap_u32(begin, begin + s2);
      } else {
        sort3(begin + s2, begin, end - 1, less);
      }
      if (!leftmost && !less(*(begin - 1), *begin)) {
        begin = partition_left(begin, end, less) + 1;
        continue;
      }
      bool already = false;
      uint32_t* pivot_pos = partition_right(begin, end, less, already);
      const int64_t l_size = pivot_pos - begin;
      const int64_t r_size = end - (pivot_pos + 1);
      if (l_size < size / 8 || r_size < size / 8) {
        if (--bad == 0) {
          heap_sort(begin, end, less);
          break;
        }
        if (l_size >= 24) {
          swap_u32(begin, begin + l_size / 4);
          swap_u32(pivot_pos - 1, pivot_pos - l_size / 4);
          if (l_size > 128) {
            swap_u32(begin + 1, begin + (l_size / 4 + 1));
            swap_u32(begin + 2, begin + (l_size / 4 + 2));
            swap_u32(pivot_pos - 2, pivot_pos - (l_size / 4 + 1));
            swap_u32(pivot_pos - 3, pivot_pos - (l_size / 4 + 2));
          }
        }
        if (r_size >= 24) {
          swap_u32(pivot_pos + 1, pivot_pos + (1 + r_size / 4));
          swap_u32(end - 1, end - r_size / 4);
          if (r_size > 128) {
            swap_u32(pivot_pos + 2, pivot_pos + (2 + r_size / 4));
            swap_u32(pivot_pos + 3, pivot_pos + (3 + r_size / 4));
            swap_u32(end - 2, end - (1 + r_size / 4));
            swap_u32(end - 3, end - (2 + r_size / 4));
          }
        }
      } else if (already && partial_ins_sort(begin, pivot_pos, less) &&
                 partial_ins_sort(pivot_pos + 1, end, less)) {
        break;
      }
      // The two partitions are disjoint and each only reads the pivot left of it, so the order
      // they are finished in does not change the result: keep the smaller one in this loop and
      // stack the larger, which bounds the stack depth by log2(n).
      if (l_size <= r_size) {
        stack[sp++] = {pivot_pos + 1, end, bad, false};
        end = pivot_pos;
      } else {
        stack[sp++] = {begin, pivot_pos, bad, leftmost};
        begin = pivot_pos + 1;
        leftmost = false;
      }
    }
  }
}

template <class L>
__host__ __device__ void split_sort(uint32_t* a, uint32_t n, uint32_t* buf, L less) {
  if (n < 2) return;
  uint32_t middle = 0;
  for (uint32_t r = 1; r < n; ++r) {
    if (less(a[r], a[middle])) {
      if (middle != 0) --middle;
    } else {
      ++middle;
      swap_u32(&a[middle], &a[r]);
    }
  }
  ++middle;
  if (middle >= n) return;
  pdq_sort(a + middle, n - middle, less);
  for (uint32_t i = 0; i < middle; ++i) buf[i] = a[i];
  uint32_t i = 0, j = middle, k = 0;
  while (i < middle && j < n) {
    if (less(a[j], buf[i])) a[k++] = a[j++];
    else a[k++] = buf[i++];
  }
  while (i < middle) a[k++] = buf[i++];
}

}  // namespace mgpu
