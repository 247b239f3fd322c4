// treecorr_amd: a device-and-host restatement of libstdc++'s std::nth_element
// (the introselect of GCC's <bits/stl_algo.h>), applied to an index array.
//
// TreeCorr's Median split and its fallback for degenerate mean splits call
// std::nth_element (src/Cell.cpp:459-473, :537-550); the order in which it leaves
// the elements fixes every later centroid sum, so the device tree build repeats
// the same comparisons and moves.  Written as plain C++ so the host test suite
// can check it against the real std::nth_element.
#pragma once

#if defined(__HIPCC__)
#define TC_HD __host__ __device__
#else
#define TC_HD
#endif

namespace tc_select {

// Compare elements by key: comp(x, y) == key[x] < key[y].
struct KeyLess {
    const double* key;
    TC_HD bool operator()(int x, int y) const { return key[x] < key[y]; }
};

TC_HD inline long lg(long n) {
    long r = 0;
    while (n > 1) { n >>= 1; ++r; }
    return r;
}

template <class Cmp>
TC_HD inline void push_heap(int* a, long hole, long top, int value, Cmp comp) {
    long parent = (hole - 1) / 2;
    while (hole > top && comp(a[parent], value)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = value;
}

template <class Cmp>
TC_HD inline void adjust_heap(int* a, long hole, long len, int value, Cmp comp) {
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (comp(a[second], a[second - 1])) second--;
        a[hole] = a[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        a[hole] = a[second - 1];
        hole = second - 1;
    }
    push_heap(a, hole, top, value, comp);
}

template <class Cmp>
TC_HD inline void make_heap(int* a, long len, Cmp comp) {
    if (len < 2) return;
    long parent = (len - 2) / 2;
    while (true) {
        int value = a[parent];
        adjust_heap(a, parent, len, value, comp);
        if (parent == 0) return;
        parent--;
    }
}

template <class Cmp>
TC_HD inline void heap_select(int* a, long middle, long last, Cmp comp) {
    make_heap(a, middle, comp);
    for (long i = middle; i < last; ++i) {
        if (comp(a[i], a[0])) {
            int value = a[i];
            a[i] = a[0];
            adjust_heap(a, 0L, middle, value, comp);
        }
    }
}

template <class Cmp>
TC_HD inline void swap_at(int* a, long i, long j) { int t = a[i]; a[i] = a[j]; a[j] = t; }

template <class Cmp>
TC_HD inline void move_median_to_first(int* a, long result, long x, long y, long z, Cmp comp) {
    if (comp(a[x], a[y])) {
        if (comp(a[y], a[z])) swap_at<Cmp>(a, result, y);
        else if (comp(a[x], a[z])) swap_at<Cmp>(a, result, z);
        else swap_at<Cmp>(a, result, x);
    } else if (comp(a[x], a[z])) swap_at<Cmp>(a, result, x);
    else if (comp(a[y], a[z])) swap_at<Cmp>(a, result, z);
    else swap_at<Cmp>(a, result, y);
}

template <class Cmp>
TC_HD inline long unguarded_partition(int* a, long first, long last, long pivot, Cmp comp) {
    while (true) {
        while (comp(a[first], a[pivot])) ++first;
        --last;
        while (comp(a[pivot], a[last])) --last;
        if (!(first < last)) return first;
        swap_at<Cmp>(a, first, last);
        ++first;
    }
}

template <class Cmp>
TC_HD inline void insertion_sort(int* a, long first, long last, Cmp comp) {
    if (first == last) return;
    for (long i = first + 1; i != last; ++i) {
        if (comp(a[i], a[first])) {
            int val = a[i];
            for (long k = i; k > first; --k) a[k] = a[k - 1];
            a[first] = val;
        } else {
            int val = a[i];
            long l = i, next = i - 1;
            while (comp(val, a[next])) {
                a[l] = a[next];
                l = next;
                --next;
            }
            a[l] = val;
        }
    }
}

// std::nth_element(a+first, a+nth, a+last, comp)
template <class Cmp>
TC_HD inline void nth_element(int* a, long first, long nth, long last, Cmp comp) {
    if (first == last || nth == last) return;
    long depth = lg(last - first) * 2;
    while (last - first > 3) {
        if (depth == 0) {
            heap_select(a + first, nth + 1 - first, last - first, comp);
            swap_at<Cmp>(a, first, nth);
            return;
        }
        --depth;
        long mid = first + (last - first) / 2;
        move_median_to_first(a, first, first + 1, mid, last - 1, comp);
        long cut = unguarded_partition(a, first + 1, last, first, comp);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    insertion_sort(a, first, last, comp);
}

}  // namespace tc_select
