"""The validator acceptance rule on the reference's own example
(period_search_pure_cpu_win/UnitTest_pure_cpu_win/InnerFunctionsTests.cpp:1059-1096):
rows from the AVX and the plain CPU builds for the same input must match."""
from validator_rule import compare, read_rows

AVX = "17.20899774  0.167425  3.307656  1.6   96   59\n"
CPU = "17.20899838  0.167423  3.307587  1.6   96   59\n"


def test_reference_example_matches():
    assert compare(AVX, CPU)


def test_rejections():
    assert not compare(AVX, "")                       # empty output
    assert not compare(AVX, AVX + CPU)                # line count differs
    near = "20.20899838  0.167423  3.307587  1.6   96   59\n"
    assert compare(AVX, near)                         # |a-b|/(a+b)/2 = 3/37.4/2 = 0.04 <= 0.1
    far = "27.20899838  0.167423  3.307587  1.6   96   59\n"
    assert not compare(AVX, far)                      # 10/44.4/2 = 0.113 > 0.1
    assert read_rows(AVX)[0][5] == 59.0
