"""AIMD root-chunk controller (system/subgraph_container.h:178-211), pinned by
the sequence worked out by hand from the reference's constants
(MAXCHUNK 1e6, DECFACTOR 8, ADDPC 0.05, MINTHRESHOLD 2000, MINSTEP 1)."""
import ctypes as C


class Aimd(C.Structure):
    _fields_ = [("chunk", C.c_uint32), ("step", C.c_uint32), ("add_step", C.c_uint32),
                ("threshold", C.c_uint32), ("recover", C.c_int32)]


def test_sequence(g2):
    L = g2.host_lib()
    L.g2h_aimd_adjust.restype = C.c_char_p
    a = Aimd()
    L.g2h_aimd_init(C.byref(a))
    assert a.chunk == 1000000
    seq = []
    for ov in [0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0]:
        tag = L.g2h_aimd_adjust(C.byref(a), ov).decode()
        seq.append((tag.split()[0], a.chunk))
    # clean at MAXCHUNK: unchanged
    assert seq[0] == ("...", 1000000)
    # overflow: /8, threshold 500000, step 6250
    assert seq[1] == ("---", 125000)
    # below the threshold the step doubles each clean iteration
    assert [c for _, c in seq[2:8]] == [131250, 143750, 168750, 218750, 318750, 518750]
    # at or above the threshold (500000) the fixed step (6250) is added
    assert seq[8] == ("+++", 525000)
    # repeated overflows divide by 8 down to 1, then stay there
    assert [c for _, c in seq[9:17]] == [65625, 8203, 1025, 128, 16, 2, 1, 1]
    assert seq[16][0] == "..."
    # recovery from 1: threshold = max(2000, 2/2) = 2000, step = max(0.05*1, 1) = 1
    assert seq[17] == ("+++", 2)
