"""Writes tests/golden/streams.npz: the first values of the seeded input streams
(mpart_amd/data.py) so that a change of generator is caught."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mpart_amd.data import Stream, banana_points, student_t_points  # noqa: E402

if __name__ == "__main__":
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "streams.npz"),
             uniform4=Stream(0).uniform(4), normal4=Stream(0).normal(4),
             banana_col0=banana_points(4)[:, 0], student_col0=student_t_points(3, 2)[:, 0])
