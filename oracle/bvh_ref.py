"""Restatement of the reference's host-side scene preparation (test
infrastructure only, pure Python, small inputs).

* ``create_bvh`` follows ``Bvh::createBvh`` (src/ray-tracing/Bvh.h:141-209)
  with its helpers ``objectBoundingBox`` (Bvh.h:94-98, eps = 1e-4 at Bvh.h:16),
  ``objectListBoundingBox`` (Bvh.h:100-114), ``boxCompare`` (Bvh.h:116-122),
  ``randomAxis`` (Bvh.h:42-45) and ``getGpuModel`` (Bvh.h:65-81).
* ``scene_lights`` follows the light-list rule of RtScene.h:87-96.

The tree depends on two third-party routines that are not in the reference:

* ``rand()`` of glibc (any 2.x; this image has 2.35).  With no ``srand`` call
  the generator is glibc's default TYPE_3 additive-feedback generator seeded
  with 1: r[0] = seed, r[i] = 16807 * r[i-1] mod (2^31 - 1) for i in 1..30,
  r[i] = r[i-31] for i in 31..33, r[i] = r[i-31] + r[i-3] (mod 2^32) for
  i >= 34, and output k is r[k+344] >> 1.
* ``std::sort`` of libstdc++ (GCC 11.4 in this image): introsort with
  2*floor(log2 n) depth limit, median-of-three pivot moved to the front
  (candidates first+1, middle, last-1), unguarded Hoare partition, loop on
  the left part and recursion on the right part above 16 elements, heap sort
  when the depth limit is exhausted, then a final insertion sort (guarded
  over the first 16 elements, unguarded after).  Its order among equal keys
  decides the tree when triangles share a minimum coordinate (meshes).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
EPS = F32(0.0001)  # Bvh.h:16  const glm::vec3 eps(0.0001)


# --------------------------------------------------------- glibc rand()
class GlibcRandom:
    """glibc random()/rand() TYPE_3 stream for a given srand seed."""

    def __init__(self, seed: int = 1):
        if seed == 0:
            seed = 1
        r = [0] * 34
        r[0] = seed & 0xFFFFFFFF
        # signed 32-bit Schrage step: 16807 * r mod 2147483647
        for i in range(1, 31):
            prev = r[i - 1]
            if prev >= 2**31:
                prev -= 2**32
            q = abs(prev) // 127773
            hi = q if prev >= 0 else -q
            lo = prev - hi * 127773
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word & 0xFFFFFFFF
        for i in range(31, 34):
            r[i] = r[i - 31]
        self._r = r
        for _ in range(310):  # outputs 0..309 are discarded (r[34..343])
            self._step()

    def _step(self) -> int:
        r = self._r
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 64:
            del r[:-34]
        return v

    def next(self) -> int:
        return self._step() >> 1


# --------------------------------------------------- libstdc++ std::sort
def _lg(n: int) -> int:
    return n.bit_length() - 1


def std_sort(a: list, less) -> None:
    """In-place libstdc++ std::sort(first, last, comp) over list ``a``."""
    n = len(a)
    if n > 1:
        _introsort_loop(a, 0, n, _lg(n) * 2, less)
        _final_insertion_sort(a, 0, n, less)


def _introsort_loop(a, first, last, depth, less):
    while last - first > 16:
        if depth == 0:
            _partial_sort(a, first, last, last, less)
            return
        depth -= 1
        cut = _unguarded_partition_pivot(a, first, last, less)
        _introsort_loop(a, cut, last, depth, less)
        last = cut


def _move_median_to_first(a, result, x, y, z, less):
    if less(a[x], a[y]):
        if less(a[y], a[z]):
            a[result], a[y] = a[y], a[result]
        elif less(a[x], a[z]):
            a[result], a[z] = a[z], a[result]
        else:
            a[result], a[x] = a[x], a[result]
    elif less(a[x], a[z]):
        a[result], a[x] = a[x], a[result]
    elif less(a[y], a[z]):
        a[result], a[z] = a[z], a[result]
    else:
        a[result], a[y] = a[y], a[result]


def _unguarded_partition(a, first, last, pivot, less):
    while True:
        while less(a[first], a[pivot]):
            first += 1
        last -= 1
        while less(a[pivot], a[last]):
            last -= 1
        if not first < last:
            return first
        a[first], a[last] = a[last], a[first]
        first += 1


def _unguarded_partition_pivot(a, first, last, less):
    mid = first + (last - first) // 2
    _move_median_to_first(a, first, first + 1, mid, last - 1, less)
    return _unguarded_partition(a, first + 1, last, first, less)


def _unguarded_linear_insert(a, last, less):
    val = a[last]
    nxt = last - 1
    while less(val, a[nxt]):
        a[last] = a[nxt]
        last = nxt
        nxt -= 1
    a[last] = val


def _insertion_sort(a, first, last, less):
    if first == last:
        return
    for i in range(first + 1, last):
        if less(a[i], a[first]):
            val = a[i]
            a[first + 1:i + 1] = a[first:i]
            a[first] = val
        else:
            _unguarded_linear_insert(a, i, less)


def _final_insertion_sort(a, first, last, less):
    if last - first > 16:
        _insertion_sort(a, first, first + 16, less)
        for i in range(first + 16, last):
            _unguarded_linear_insert(a, i, less)
    else:
        _insertion_sort(a, first, last, less)


def _push_heap(a, first, hole, top, value, less):
    parent = (hole - 1) // 2
    while hole > top and less(a[first + parent], value):
        a[first + hole] = a[first + parent]
        hole = parent
        parent = (hole - 1) // 2
    a[first + hole] = value


def _adjust_heap(a, first, hole, length, value, less):
    top = hole
    second = hole
    while second < (length - 1) // 2:
        second = 2 * (second + 1)
        if less(a[first + second], a[first + second - 1]):
            second -= 1
        a[first + hole] = a[first + second]
        hole = second
    if (length & 1) == 0 and second == (length - 2) // 2:
        second = 2 * (second + 1)
        a[first + hole] = a[first + second - 1]
        hole = second - 1
    _push_heap(a, first, hole, top, value, less)


def _make_heap(a, first, last, less):
    length = last - first
    if length < 2:
        return
    parent = (length - 2) // 2
    while True:
        _adjust_heap(a, first, parent, length, a[first + parent], less)
        if parent == 0:
            return
        parent -= 1


def _pop_heap(a, first, last, result, less):
    value = a[result]
    a[result] = a[first]
    _adjust_heap(a, first, 0, last - first, value, less)


def _partial_sort(a, first, middle, last, less):
    _make_heap(a, first, middle, less)
    for i in range(middle, last):
        if less(a[i], a[first]):
            _pop_heap(a, first, middle, i, less)
    while middle - first > 1:
        middle -= 1
        _pop_heap(a, first, middle, middle, less)


# ------------------------------------------------------------ the builder
def _gmin(x, y):
    return y if y < x else x


def _gmax(x, y):
    return y if x < y else x


def object_bounding_box(v0, v1, v2):
    """objectBoundingBox, Bvh.h:94-98 (float32, glm min/max semantics)."""
    mn = [F32(_gmin(_gmin(F32(v0[k]), F32(v1[k])), F32(v2[k])) - EPS) for k in range(3)]
    mx = [F32(_gmax(_gmax(F32(v0[k]), F32(v1[k])), F32(v2[k])) + EPS) for k in range(3)]
    return mn, mx


def create_bvh(triangles: np.ndarray, rng: GlibcRandom | None = None) -> np.ndarray:
    """Bvh::createBvh over a TRI_DT array; returns a NODE_DT array (2n-1)."""
    from .oracle import NODE_DT

    n = len(triangles)
    if n == 0:
        raise ValueError("createBvh of an empty list is undefined (Bvh.h:77)")
    rng = rng or GlibcRandom(1)
    boxes = [object_bounding_box(t["v0"], t["v1"], t["v2"]) for t in triangles]
    # Object0 {index, triangle}: the comparator only looks at the box min.
    objects = [(i, boxes[i][0]) for i in range(n)]

    intermediate = []  # (index, box, left, right, objects)
    counter = 0
    stack = [(counter, objects)]
    counter += 1
    while stack:
        index, objs = stack.pop()
        # objectListBoundingBox, Bvh.h:100-114
        bmn, bmx = None, None
        for (i, _) in objs:
            tmn, tmx = boxes[i]
            if bmn is None:
                bmn, bmx = list(tmn), list(tmx)
            else:
                bmn = [_gmin(bmn[k], tmn[k]) for k in range(3)]
                bmx = [_gmax(bmx[k], tmx[k]) for k in range(3)]
        axis = rng.next() % 3
        objs = list(objs)
        std_sort(objs, lambda a, b, ax=axis: a[1][ax] < b[1][ax])
        span = len(objs)
        if span <= 1:
            intermediate.append((index, bmn, bmx, -1, -1, objs))
            continue
        mid = span // 2
        left_index = counter
        counter += 1
        stack.append((left_index, objs[:mid]))
        right_index = counter
        counter += 1
        stack.append((right_index, objs[mid:]))
        intermediate.append((index, bmn, bmx, left_index, right_index, objs))
    intermediate.sort(key=lambda r: r[0])

    out = np.zeros(len(intermediate), NODE_DT)
    for k, (index, bmn, bmx, left, right, objs) in enumerate(intermediate):
        out[k]["min"] = bmn
        out[k]["max"] = bmx
        out[k]["left"] = left
        out[k]["right"] = right
        out[k]["object"] = objs[0][0] if (left == -1 and right == -1) else -1
    return out


def scene_lights(triangles: np.ndarray, materials: np.ndarray) -> np.ndarray:
    """RtScene.h:87-96: Light{i, length(cross(v0, v1)) * 0.5f} per light triangle."""
    from .oracle import LIGHT_DT

    out = []
    for i, t in enumerate(triangles):
        if int(materials[int(t["materialIndex"])]["type"]) == 0:
            a = [F32(x) for x in t["v0"]]
            b = [F32(x) for x in t["v1"]]
            c = [F32(a[1] * b[2] - b[1] * a[2]), F32(a[2] * b[0] - b[2] * a[0]), F32(a[0] * b[1] - b[0] * a[1])]
            d = F32(F32(F32(c[0] * c[0]) + F32(c[1] * c[1])) + F32(c[2] * c[2]))
            area = F32(F32(np.sqrt(d)) * F32(0.5))
            out.append((i, area))
    arr = np.zeros(len(out), LIGHT_DT)
    for k, (i, a) in enumerate(out):
        arr[k]["triangleIndex"] = i
        arr[k]["area"] = a
    return arr
