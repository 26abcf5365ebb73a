"""The checkpoint Header reader on a Header laid out as
LBM::write_checkpoint_file writes it (Source/LBM.cpp:1711-1738: title,
finest level, steps, dts and times each followed by a space, then every
level's BoxArray as BoxArray::writeOn prints it); no GPU needed."""
import pytest

HEADER = """Checkpoint file for LBM
1
40 80 \n0.5 0.25 \n20 20.000000000000004 \n(2 0
((0,0,0) (15,31,31) (0,0,0))
((16,0,0) (31,31,31) (0,0,0))
)
(1 0
((8,8,8) (23,39,23) (0,0,0))
)
"""


def test_header_fields_and_box_arrays(engine, tmp_path):
    from marbles_amd import checkpoint

    (tmp_path / "Header").write_text(HEADER)
    info = checkpoint.checkpoint_info(str(tmp_path))
    assert info["nlevels"] == 2
    assert info["steps"] == [40, 80]
    assert info["dts"] == [0.5, 0.25]
    assert info["times"] == [20.0, 20.000000000000004]
    assert info["boxes"][0] == [((0, 0, 0), (15, 31, 31)), ((16, 0, 0), (31, 31, 31))]
    assert info["boxes"][1] == [((8, 8, 8), (23, 39, 23))]


def test_missing_level_is_an_error(engine, tmp_path):
    from marbles_amd import checkpoint
    from marbles_amd._lib import MrbError

    (tmp_path / "Header").write_text(HEADER.replace("Checkpoint file for LBM\n1\n", "Checkpoint file for LBM\n2\n"))
    with pytest.raises(MrbError):
        checkpoint.checkpoint_info(str(tmp_path))
