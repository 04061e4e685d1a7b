"""Literal CPU restatement of futhark/parser.fut (test infrastructure, not product code).

Follows ``mk_parser``:
  * keys over ``[start] ++ tokens ++ [end]`` (``get_key``, parser.fut:117-122), looked up in the
    open-addressing table with the FNV-style ``hash`` and the ``max_iters`` probe loop
    (parser.fut:110-151); any miss rejects (``to_keys``, :235-240);
  * ``segmented_copy`` of the stack and production spans (:216-233);
  * ``depths``: +-1 scan, reject on a negative prefix or a non-zero total (:153-166);
  * ``mk_tree``/``find_previous``: the implicit min-tree and its "previous <=" walk (:72-108);
  * ``brackets_matches`` (:174-184), ``parents`` (:286-296), ``terminal_offsets`` and the final
    scatter of terminal nodes (:305-332).
The segmented copy uses ``repl_segm_iota`` from github.com/diku-dk/segmented (unpinned; restated
from its definition: segment index and in-segment offset of every output element).
"""
from __future__ import annotations

import numpy as np

HIGHEST = np.iinfo(np.int64).max
_M64 = (1 << 64) - 1


def fnv_hash(keys: np.ndarray) -> np.ndarray:
    h = np.full(keys.shape[0], 14695981039346656037, dtype=np.uint64)
    prime = np.uint64(1099511628211)
    for j in range(keys.shape[1]):
        h = (h ^ keys[:, j].astype(np.uint64)) * prime
    return h


def mk_tree(arr: np.ndarray) -> np.ndarray:
    """parser.fut:72-90 with op = min, ne = i64.highest."""
    n = arr.shape[0]
    temp = int(n).bit_length()
    h = temp if bin(n).count("1") == 1 else temp + 1
    tree_size = (1 << h) - 1
    offset = (1 << (h - 1)) - 1
    tree = np.full(tree_size, HIGHEST, dtype=np.int64)
    tree[offset:offset + n] = arr
    level = h - 2
    cur = tree[offset:].copy()
    while level >= 0:
        cur = np.minimum(cur[0::2], cur[1::2])
        off = (1 << level) - 1
        tree[off:off + cur.shape[0]] = cur
        level -= 1
    return tree


def find_previous_le(tree: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """``find_previous (<=)`` for many indices at once (parser.fut:92-108)."""
    n = tree.shape[0]
    h = int(n).bit_length()
    offset = (1 << (h - 1)) - 1
    cur = offset + idx.astype(np.int64)
    v = tree[cur]

    def sibling(i):
        return i - (i % 2 == 0) + (i % 2 == 1)

    while True:
        sib = np.clip(sibling(cur), 0, n - 1)
        asc = (cur != 0) & ((cur % 2 == 1) | ~(tree[sib] <= v))
        if not asc.any():
            break
        cur = np.where(asc, (cur - 1) // 2, cur)
    found = cur != 0
    i = np.where(found, sibling(cur), 0)
    while True:
        go = found & (i < offset)
        if not go.any():
            break
        right = 2 * i + 2
        step = 2 * i + 1 + (tree[np.clip(right, 0, n - 1)] <= v)
        i = np.where(go, step, i)
    return np.where(found, i - offset, -1)


class ParserFut:
    def __init__(self, parser_tables):
        pt = parser_tables
        self.pt = pt
        self.q, self.k = pt.q, pt.k
        self.size = pt.hash_used.shape[0]
        self.top = np.int64(1 << (pt.bracket_bits - 1))

    def lookup(self, keys: np.ndarray) -> np.ndarray:
        """Vectorised ``lookup`` (parser.fut:130-142); returns spans [m, 4] or -1s."""
        pt = self.pt
        m = keys.shape[0]
        out = np.full((m, 4), -1, dtype=np.int64)
        if self.size == 0 or m == 0:
            return out
        h = fnv_hash(keys) % np.uint64(self.size)
        h = h.astype(np.int64)
        is_found = np.zeros(m, dtype=bool)
        i = 0
        active = np.ones(m, dtype=bool)
        while True:
            active = is_found | (i < pt.max_iters)
            if not active.any():
                break
            valid = active & pt.hash_used[h] & np.all(pt.hash_keys[h] == keys, axis=1)
            out = np.where(valid[:, None], pt.hash_vals[h], out)
            is_found = np.where(active, valid, is_found)
            h = np.where(active, (h + 1) % self.size, h)
            i += 1
            if i > pt.max_iters + self.size + 2:
                break
        return out

    def keys(self, terms: np.ndarray) -> np.ndarray:
        pt = self.pt
        arr = np.concatenate([[pt.start_terminal], terms.astype(np.int64), [pt.end_terminal]])
        n = arr.shape[0]
        q, k = self.q, self.k
        cols = []
        for j in range(q + k):
            pos = np.arange(n) + j - q
            ok = (pos >= 0) & (pos < n)
            cols.append(np.where(ok, arr[np.clip(pos, 0, n - 1)], pt.empty_terminal))
        return np.stack(cols, axis=1) if cols else np.zeros((n, 0), dtype=np.int64)

    @staticmethod
    def segmented_copy(arr: np.ndarray, starts: np.ndarray, ends: np.ndarray) -> np.ndarray:
        shape = ends - starts
        total = int(shape.sum())
        if total == 0:
            return np.zeros(0, dtype=arr.dtype)
        seg = np.repeat(np.arange(shape.shape[0]), shape)
        first = np.concatenate([[0], np.cumsum(shape)[:-1]])
        inner = np.arange(total) - first[seg]
        return arr[starts[seg] + inner]

    def brackets_matches(self, br: np.ndarray) -> bool:
        n = br.shape[0]
        left = (br & self.top) != 0
        scan = np.cumsum(np.where(left, 1, -1))
        if (scan < 0).any() or (n != 0 and scan[-1] != 0):
            return False
        if n == 0:
            return True
        ds = scan - left.astype(np.int64)
        tree = mk_tree(ds)
        rights = np.nonzero(~left)[0]
        j = find_previous_le(tree, rights)
        if (j < 0).any():
            return False
        mask = ~self.top
        return bool(np.all((br[rights] & mask) == (br[j] & mask)))

    def pre_productions_int(self, terms: np.ndarray):
        spans = self.lookup(self.keys(terms))
        if (spans == -1).any():
            return None
        stacks = self.segmented_copy(self.pt.flat_stacks, spans[:, 0], spans[:, 1])
        if not self.brackets_matches(stacks):
            return None
        return self.segmented_copy(self.pt.flat_productions, spans[:, 2], spans[:, 3])

    def parents(self, prods: np.ndarray) -> np.ndarray:
        n = prods.shape[0]
        if n == 0:
            return np.zeros(0, dtype=np.int64)
        a = self.pt.arities[prods].astype(np.int64) - 1
        s = np.concatenate([[0], np.cumsum(a)[:-1]])
        tree = mk_tree(s)
        par = find_previous_le(tree, np.arange(n))
        par[0] = 0
        return par

    def parse_int(self, kinds, starts, ends):
        """Returns None on rejection, else (parent, is_terminal, id, start, end) int64 arrays."""
        prods = self.pre_productions_int(np.asarray(kinds, dtype=np.int64))
        if prods is None:
            return None
        n = prods.shape[0]
        par = self.parents(prods)
        p2t = self.pt.production_to_terminal[prods]
        is_t = p2t >= 0
        tidx = np.nonzero(is_t)[0]
        node_id = prods.astype(np.int64).copy()
        st = np.zeros(n, dtype=np.int64)
        en = np.zeros(n, dtype=np.int64)
        if tidx.shape[0] == len(starts):        # safe_zip (parser.fut:300-303)
            node_id[tidx] = p2t[tidx]
            st[tidx] = np.asarray(starts, dtype=np.int64)
            en[tidx] = np.asarray(ends, dtype=np.int64)
        else:
            is_t = np.zeros(n, dtype=bool)
        return par, is_t, node_id, st, en
