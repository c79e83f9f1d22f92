"""The drivers' file formats through the C ABI (host code only, no GPU).

Known answers: the subset and config readers follow `infile >> int` and
`getline` (Tree_u_NC.cpp:338-371); the embedding CSV is Eigen's
IOFormat(StreamPrecision, DontAlignCols, ", ", "\\n") (Tree_u_NC.cpp:63,
898-909), i.e. every coefficient written by an ostream at its default
precision 6 in general notation, which prints 1/3 as 0.333333, 1e-5 as
1e-05, 123456789 as 1.23457e+08 and -0.0 as -0."""
import numpy as np
import pytest


def test_embedding_csv_is_eigen_stream_precision_format(capi, tmp_path):
    U = np.array([[1.0 / 3.0, -1e-5, 123456789.0], [2.0, -0.0, 0.1], [1e-300, 1e22, -7.25]])
    p = tmp_path / "U.csv"
    capi.write_embedding_csv(str(p), U)
    assert p.read_text() == ("0.333333, -1e-05, 1.23457e+08\n"
                             "2, -0, 0.1\n"
                             "1e-300, 1e+22, -7.25")


def test_embedding_csv_empty_and_single(capi, tmp_path):
    p = tmp_path / "e.csv"
    capi.write_embedding_csv(str(p), np.zeros((0, 3)))
    assert p.read_text() == ""
    capi.write_embedding_csv(str(p), np.array([[5.0]]))
    assert p.read_text() == "5"


def test_subset_reader_stops_at_first_non_integer(capi, tmp_path):
    p = tmp_path / "s.txt"
    p.write_text("12\n7 9\n\n4\n")
    assert capi.read_subset_file(str(p)).tolist() == [12, 7, 9, 4]
    p.write_text("3 5 x 8\n")
    assert capi.read_subset_file(str(p)).tolist() == [3, 5]
    p.write_text("")
    assert capi.read_subset_file(str(p)).tolist() == []
    with pytest.raises(capi.TreeSVDError):
        capi.read_subset_file(str(tmp_path / "missing.txt"))


def test_config_lines_follow_getline(capi, tmp_path):
    p = tmp_path / "config.txt"
    p.write_text("sub.txt\nsnap1.txt\nsnap2.txt\n")
    assert capi.read_lines(str(p)) == ["sub.txt", "snap1.txt", "snap2.txt"]
    p.write_text("a\n\nb")                   # empty middle line kept, last line without newline counted
    assert capi.read_lines(str(p)) == ["a", "", "b"]
    p.write_text("")
    assert capi.read_lines(str(p)) == []
