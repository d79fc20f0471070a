"""Ordered subsets of projection angles, Python/tigre/utilities/order_subsets.py:31-96."""
from __future__ import annotations

import numpy as np


def _angular_distance_index(angles):
    """Greedy ordering that keeps consecutive views ~90 degrees apart (order_subsets.py:6-29)."""
    alpha = np.asarray(angles)[:, 0]
    order = list(np.argsort(alpha))
    chosen = [order.pop(0)]
    while order:
        last = alpha[chosen[-1]]
        k = int(np.argmin(np.abs(np.abs(alpha[order] - last) - np.pi / 2)))
        chosen.append(order.pop(k))
    return np.array(chosen)


def order_subsets(angles, blocksize, mode):
    """Return (angle blocks, index blocks) for the given strategy:
    None/'ordered', 'random' (shuffled blocks), 'random2' (shuffled views),
    'angularDistance'."""
    angles = np.asarray(angles)
    n = angles.shape[0]
    if blocksize is None or blocksize == 1:
        index = np.arange(n)
        if mode is None or mode == "ordered":
            return list(angles), list(index)
        if mode in ("random", "random2"):
            np.random.shuffle(index)
            return list(angles[index]), list(index)
        if mode == "angularDistance":
            index = _angular_distance_index(angles)
            return list(np.array(angles[index], dtype=np.float32)), list(index)
        raise NameError("mode string not recognized: " + str(mode))
    count = (n + blocksize - 1) // blocksize

    def split(order):
        idx = [order[i * blocksize:(i + 1) * blocksize] for i in range(count)]
        return [angles[i] for i in idx], idx

    if mode == "random2":
        order = np.arange(n)
        np.random.shuffle(order)
        return split(order)
    if mode == "angularDistance":
        return split(_angular_distance_index(angles))
    blocks, idx = split(np.arange(n, dtype=np.int32))
    if mode is None or mode == "ordered":
        return blocks, idx
    if mode == "random":
        perm = np.arange(len(idx))
        np.random.shuffle(perm)
        return [blocks[i] for i in perm], [idx[i] for i in perm]
    raise NameError("mode string not recognized: " + str(mode))
