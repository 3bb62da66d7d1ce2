"""Pin of the selective-GCA collector oracle (oracle/gca_select.py): a single
large coefficient value at one node marks exactly the cells whose wedges hold
that node; a smooth coefficient marks nothing.  CPU only."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import gca_select as GS  # noqa: E402


def test_single_node_spike_marks_its_cells():
    n, nr = 9, 9
    k = np.ones((10, n, n, nr))
    d0, x0, y0, r0 = 3, 4, 5, 4
    k[d0, x0, y0, r0] = 100.0
    flags = GS.collect(k, n, nr, 0, n, nr)
    expected = np.zeros_like(flags)
    for x in range(n - 1):
        for y in range(n - 1):
            for r in range(nr - 1):
                for w in range(2):
                    if any((x + ox, y + oy, r + orr) == (x0, y0, r0) for ox, oy, orr in GS.WEDGE_NODES[w]):
                        expected[d0, x, y, r] = 1.0
    assert np.array_equal(flags, expected)
    # an interior node belongs to (a wedge of) each of the 8 hex cells around it
    assert int(expected.sum()) == 8
    # mapped two levels up, the marks land on the ancestors (index >> 2)
    up = GS.collect(k, n, nr, 2, 3, 3)
    anc = np.zeros_like(up)
    for d, x, y, r in zip(*np.nonzero(expected)):
        anc[d, x >> 2, y >> 2, r >> 2] = 1.0
    assert np.array_equal(up, anc)


def test_gradient_threshold():
    n, nr = 5, 5
    # k linear in the reference coordinate xi of every cell: grad_ref = (c, 0, 0) per wedge (up to the wedge's
    # orientation), so c = 9 stays below the threshold and c = 11 crosses it
    xi = np.arange(n)[None, :, None, None] * np.ones((10, n, n, nr))
    for c, marked in ((9.0, False), (11.0, True)):
        flags = GS.collect(c * xi, n, nr, 0, n, nr)
        assert bool(flags.any()) == marked
