// introsort.h -- the permutation libstdc++'s std::sort produces, computed on
// the GPU (and on the host, for tests).
//
// The reference orders cluster-path candidates with std::sort on a
// "covariance descending" comparator (sugarbox_indexed_neighbour_lookup.h:100,
// entry_t at :81-105).  The sort is not stable and exact covariance ties are
// common, so which tied candidates land in the first max_neighbours slots is
// decided by the library's algorithm.  To reproduce the neighbour sets the
// same steps have to run: libstdc++'s introsort as published in GCC's
// bits/stl_algo.h / bits/stl_heap.h (GCC 11):
//   - introsort loop while a range holds more than 16 elements, depth limit
//     2*floor(log2 n); pivot = median of (first+1, mid, last-1) moved to
//     first, then an unguarded Hoare partition of [first+1, last);
//     recursion on the right part, iteration on the left;
//   - heapsort (make_heap + sort_heap) once the depth limit is spent;
//   - a final insertion sort: guarded on the first 16 elements, unguarded on
//     the rest.
// This file is an independent implementation of those steps over a
// structure-of-arrays store; its output is checked against std::sort itself
// on tie-heavy inputs by the test suite.
#pragma once

#if defined(__HIPCC__)
#define HPGLX_HD __host__ __device__
#else
#define HPGLX_HD
#endif

namespace hpglx {

// Store concept: key(i) -> double, swap(i, j), get(i) -> Item, put(i, Item);
// "less" is key(a) > key(b) (larger covariance first).
template <typename Store>
struct Introsort {
    typedef typename Store::Item Item;
    Store& s;
    HPGLX_HD explicit Introsort(Store& st) : s(st) {}

    HPGLX_HD bool lt_ii(int a, int b) const { return s.key(a) > s.key(b); }
    HPGLX_HD static bool lt_vv(const Item& a, const Item& b) { return a.key > b.key; }

    // The three keys are read once, before any swap (lt(x, y) is key(x) > key(y)).
    HPGLX_HD void move_median_to_first(int result, int a, int b, int c) {
        const auto ka = s.key(a), kb = s.key(b), kc = s.key(c);
        if (ka > kb) {
            if (kb > kc)
                s.swap(result, b);
            else if (ka > kc)
                s.swap(result, c);
            else
                s.swap(result, a);
        } else if (ka > kc)
            s.swap(result, a);
        else if (kb > kc)
            s.swap(result, c);
        else
            s.swap(result, b);
    }

    // The library's loop scans from the left to an element not less than the
    // pivot, then from the right to one not greater, swaps the two while they
    // have not crossed, and returns the left position once they have.  Both
    // scans read an array that no scan changes, so they can move in the same
    // steps: each step reads both ends, moves the scanner(s) that have not
    // stopped, and once both have stopped swaps (or returns).  The positions
    // read, the swaps and the result are the library's; a GPU lane takes one
    // step per LDS latency instead of one per element and scan, and lanes of
    // a wave run one loop instead of three nested ones.  The pivot never
    // moves meanwhile (it sits before first), so its key is read once.
    HPGLX_HD int unguarded_partition(int first, int last, int pivot) {
        const auto kp = s.key(pivot);
        --last;
        while (true) {
            const Item a = s.get(first), b = s.get(last);
            const bool left_on = a.key > kp;   // a is less than the pivot
            const bool right_on = kp > b.key;  // the pivot is less than b
            if (!left_on && !right_on) {
                if (!(first < last)) return first;
                s.put(first, b);
                s.put(last, a);
                ++first;
                --last;
            } else {
                first += left_on ? 1 : 0;
                last -= right_on ? 1 : 0;
            }
        }
    }

    // Heap helpers on [first, first + len).
    HPGLX_HD void push_heap(int first, int hole, int top, Item value) {
        int parent = (hole - 1) / 2;
        while (hole > top && lt_vv(s.get(first + parent), value)) {
            s.put(first + hole, s.get(first + parent));
            hole = parent;
            parent = (hole - 1) / 2;
        }
        s.put(first + hole, value);
    }

    HPGLX_HD void adjust_heap(int first, int hole, int len, Item value) {
        const int top = hole;
        int child = hole;
        while (child < (len - 1) / 2) {
            child = 2 * (child + 1);
            if (lt_ii(first + child, first + (child - 1))) child--;
            s.put(first + hole, s.get(first + child));
            hole = child;
        }
        if ((len & 1) == 0 && child == (len - 2) / 2) {
            child = 2 * (child + 1);
            s.put(first + hole, s.get(first + (child - 1)));
            hole = child - 1;
        }
        push_heap(first, hole, top, value);
    }

    HPGLX_HD void make_heap(int first, int last) {
        const int len = last - first;
        if (len < 2) return;
        int parent = (len - 2) / 2;
        while (true) {
            Item v = s.get(first + parent);
            adjust_heap(first, parent, len, v);
            if (parent == 0) return;
            parent--;
        }
    }

    HPGLX_HD void pop_heap(int first, int last, int result) {
        Item v = s.get(result);
        s.put(result, s.get(first));
        adjust_heap(first, 0, last - first, v);
    }

    HPGLX_HD void heap_sort(int first, int last) {
        // heap_select with middle == last reduces to make_heap.
        make_heap(first, last);
        while (last - first > 1) {
            --last;
            pop_heap(first, last, last);
        }
    }

    HPGLX_HD void unguarded_linear_insert(int last) {
        Item v = s.get(last);
        int next = last - 1;
        while (lt_vv(v, s.get(next))) {
            s.put(last, s.get(next));
            last = next;
            --next;
        }
        s.put(last, v);
    }

    HPGLX_HD void insertion_sort(int first, int last) {
        if (first == last) return;
        for (int i = first + 1; i != last; ++i) {
            if (lt_ii(i, first)) {
                Item v = s.get(i);
                for (int k = i; k > first; --k) s.put(k, s.get(k - 1));
                s.put(first, v);
            } else {
                unguarded_linear_insert(i);
            }
        }
    }

    HPGLX_HD static int lg(int n) {
        int r = 0;
        while (n > 1) {
            n >>= 1;
            ++r;
        }
        return r;
    }

    // Recursion of the right-hand part is replaced by an explicit stack of
    // (first, last, depth) ranges, visited in the same order.
    //
    // With kneed < n only positions [0, kneed) are wanted: a right-hand part
    // that starts at or after kneed is left unsorted.  That changes nothing
    // before it: partitioning never moves elements across a cut, the
    // left-hand part continues with the same depth, and the final insertion
    // sort cannot move an element leftwards across a cut (an element is
    // never "less" than one on the left of a cut), so it only has to run up
    // to the first unsorted part.
    HPGLX_HD void sort(int n, int kneed = 0x7fffffff) {
        if (n <= 0) return;
        int lim = n;  // [0, lim) ends up fully sorted
        const int kThreshold = 16;
        struct Range { int first, last, depth; };
        Range stack[64];
        int sp = 0;
        // introsort_loop(first, last, depth): while (len > 16) {...; recurse(cut, last); last = cut;}
        // The recursive call on [cut, last) must complete before the loop
        // continues on [first, cut); a LIFO stack of pending left parts does that.
        int first = 0, last = n, depth = 2 * lg(n);
        while (true) {
            while (last - first > kThreshold) {
                if (depth == 0) {
                    heap_sort(first, last);
                    last = first;  // nothing left in this range
                    break;
                }
                --depth;
                const int mid = first + (last - first) / 2;
                move_median_to_first(first, first + 1, mid, last - 1);
                const int cut = unguarded_partition(first + 1, last, first);
                if (cut >= kneed) {
                    // [cut, last) is not wanted: go on with [first, cut)
                    lim = cut < lim ? cut : lim;
                    last = cut;
                    continue;
                }
                // Defer [first, cut) with the current depth, descend into [cut, last).
                stack[sp].first = first;
                stack[sp].last = cut;
                stack[sp].depth = depth;
                ++sp;
                first = cut;
            }
            if (sp == 0) break;
            --sp;
            first = stack[sp].first;
            last = stack[sp].last;
            depth = stack[sp].depth;
        }
        // final_insertion_sort, over [0, lim): the same steps as over [0, n)
        // for every element before lim
        if (lim > kThreshold) {
            insertion_sort(0, kThreshold);
            for (int i = kThreshold; i != lim; ++i) unguarded_linear_insert(i);
        } else {
            insertion_sort(0, lim);
        }
    }
};

}  // namespace hpglx
