"""The CPU restatement of the device generator (tests/philox_ref.py) against
the published Philox-4x32-10 known-answer vectors (Random123 kat_vectors,
the SC'11 paper's generator), and the distribution of its Box-Muller
output."""
import numpy as np

from philox_ref import gaussian_stream, philox4x32_10

KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answer_vectors():
    ctr = np.array([k[0] for k in KAT], dtype=np.uint64)
    key = np.array([k[1] for k in KAT], dtype=np.uint64)
    out = philox4x32_10(ctr, key)
    for row, (_, _, want) in zip(out, KAT):
        assert tuple(int(x) for x in row) == want


def test_gaussian_stream_distribution_and_offsets():
    z = gaussian_stream(7, 0, 400000)
    assert abs(z.mean()) < 5e-3 and abs(z.std() - 1.0) < 5e-3
    assert abs(np.mean(z ** 4) - 3.0) < 0.05
    # a draw from position 2k continues the stream from 0
    assert np.array_equal(gaussian_stream(7, 1000, 50), z[1000:1050])
