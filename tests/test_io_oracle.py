"""The state-file oracle against hand-worked cases of the reference's
reader (aux.c:131-156) and writer (write.c:289-299)."""
import numpy as np

from oracle import io_oracle as IO


def block(n, stlen):
    # state k = (k, -k, 10k, ...): every value names its state
    return np.array([[k * (1 + j) * (-1) ** j for j in range(stlen)] for k in range(n)], dtype=np.float64)


def test_writer_layout():
    s = block(4, 3)
    raw = IO.write_psi_dat(s, 3)
    assert raw == s.astype("<f8").tobytes()
    raw = IO.write_psi_dat(s, 3, idx=[2, 0])
    assert np.array_equal(np.frombuffer(raw, "<f8"), np.concatenate([s[2], s[0]]))
    raw = IO.write_psi_dat(s, 3, idx=[3, 1], n_every=2)
    assert np.array_equal(np.frombuffer(raw, "<f8"), np.concatenate([np.zeros(3), s[3], np.zeros(3), s[1]]))
    assert IO.write_psi_dat(s, 3, idx=[]) == b""


def test_reader_by_hand():
    s = block(6, 2)
    raw = s.astype("<f8").tobytes()
    # n_every = 2 from 0: seek 0, then skip one and read one, three times: states 1, 3, 5
    assert np.array_equal(IO.read_psi_dat(raw, 0, 6, 2, 2), s[[1, 3, 5]])
    # from 1 to 5: i = 1 -> state 2, i = 3 -> state 4
    assert np.array_equal(IO.read_psi_dat(raw, 1, 5, 2, 2), s[[2, 4]])
    # stride 1 reads start .. end-1
    assert np.array_equal(IO.read_psi_dat(raw, 2, 4, 2, 1), s[[2, 3]])
    # a selection past the data ends the list
    assert np.array_equal(IO.read_psi_dat(raw, 3, 9, 2, 2), s[[4]])


def test_round_trip_and_placement():
    s = block(5, 4)
    raw = IO.write_psi_dat(s, 4, n_every=2)
    assert len(raw) == 5 * 2 * 4 * 8
    back = IO.read_psi_dat(raw, 0, 10, 4, 2)
    assert np.array_equal(back, s)
    blocks = IO.start_blocks(back, 3, 2, 4).reshape(3, 2, 4)
    assert np.array_equal(blocks[:, 0], s[:3]) and not blocks[:, 1].any()
