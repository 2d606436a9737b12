"""Writes tests/golden/fixtures.json: the literal known-answer cases that the
reference's own test suite asserts (inputs and expected outputs transcribed
from the assertions; each entry names the test file and lines).  Run it from
the repository root: ``python tests/golden/make_fixtures.py``.

Matrices are given densely (column-major lists of columns); ``None`` stands
for NA.  Expected values that the reference tests state as properties rather
than numbers are stored as property records checked by the tests.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def cols(*cs):
    return [list(map(float, c)) for c in cs]


fixtures = {
    "topk_na_padding": {
        "source": "tests/testthat/test-sparseKNN.R:182-194",
        "D": [[1.0, None, "Inf"]],  # one column
        "k": 3, "decreasing": False, "self_row": [0],
        "expect_idx": [[1, None, None]],
        "expect_dist": [[1.0, None, None]],
    },
    "topk_validation": {
        "source": "tests/testthat/test-sparseKNN.R:196-206",
        "cases": [
            {"self_row": [1, 99], "k": 2, "error": "valid 1-based rows"},
            {"self_row": [1, -1], "k": 2, "error": "valid 1-based rows"},
            {"self_row": [0, 0], "k": 0, "error": "k must be"},
        ],
    },
    "ties_lower_index": {
        "source": "tests/testthat/test-sparseKNN.R:160-168",
        "X": [[1, 0, 1, 0]] * 6, "method": "euclidean", "k": 3,
        "row": 4, "expect_idx_row": [1, 2, 3],
    },
    "massive_jaccard_ties": {
        "source": "tests/testthat/test-sparseKNNCross.R:93-104",
        "nrow": 50, "ncol_x": 40, "ncol_y": 25, "ones_rows": 20, "k": 6,
        "block_sizes": [1, 3, 8, 40],
        "expect_forward_row": [1, 2, 3, 4, 5, 6], "expect_reverse_row": [1, 2, 3, 4, 5, 6],
        "expect_dist": 0.0,
    },
    "signed_cancel_not_empty": {
        "source": "tests/testthat/test-sparseKNN.R:415-433",
        "X": cols([1, -1, 0, 0], [0, 0, 0, 0], [1, 0, 1, 0], [0, 1, 0, 1]),
        "k": 2, "methods": ["euclidean", "cosine"],
        "expect": {"row1_complete": True, "row2_all_na": True, "never_returns": 2,
                   "some_row_returns": 1},
    },
    "cross_degenerate": {
        "source": "tests/testthat/test-sparseKNN.R:435-465",
        "clean": cols([1, 0, 1, 0], [0, 1, 0, 1], [1, 1, 0, 0]),
        "holed": cols([1, 0, 0, 1], [0, 0, 0, 0], [0, 1, 1, 0]),
        "k": 2,
        "expect": {
            "a_holed_vs_clean": {"na_rows": [2], "complete_rows": [1, 3]},
            "b_clean_vs_holed": {"never_returns": 2, "no_na": True},
            "c_holed_vs_holed": {"na_rows": [2], "never_returns": 2},
            "d_clamp": {"k": 3, "expect_k": 2, "warns": "exceeds"},
        },
    },
    "pearson_constant_self": {
        "source": "tests/testthat/test-sparseKNN.R:254-274",
        "X": cols([1, 1, 1, 1, 1], [1, 2, 3, 4, 5]),
        "k": 2,
        "expect_sim_self_value": 1.0,
        "expect_dist_idx_col1": [1, 2], "expect_dist_dist_col1": [0.0, 0.0],
    },
    "cosine_extremes": {
        "source": "tests/testthat/test-sparseDist.R:301-328",
        "X": cols([1e200, 0], [1e200, 0], [1e-200, 0], [1e-200, 0]),
        "expect_sim": [[1, 2, 1.0], [3, 4, 1.0], [1, 3, 1.0]],
    },
    "cosine_antiparallel": {
        "source": "tests/testthat/test-sparseDist.R:260-263",
        "X": cols([1, 2], [-1, -2]),
        "expect_sim": [[1, 2, -1.0]],
    },
    "empty_conventions": {
        "source": "tests/testthat/test-sparseDist.R:357-378",
        "X": cols([0, 0], [0, 0], [1, 2]),
        "expect_sim": [[1, 2, 1.0], [1, 3, 0.0], [1, 1, 1.0], [2, 2, 1.0], [3, 3, 1.0]],
        "expect_cosine_cross_dist": [[1, 1, 0.0], [2, 2, 0.0]],
        "cosine_cross_dist_33_below": 1e-12,
    },
    "snn_self_listed": {
        "source": "tests/testthat/test-sparseSNN.R:55-57",
        "idx": [[1, 1], [2, 2], [3, 3]], "include_self": True,
        "expect_w_col2": [1.0, 1.0, 1.0],
    },
    "snn_disjoint": {
        "source": "tests/testthat/test-sparseSNN.R:60-69",
        "idx": [[2, 3], [1, 4], [1, 3], [2, 4]], "include_self": False,
        "expect_w": [[0.0, 0.0], [0.0, 0.0], [0.0, 1.0], [0.0, 1.0]],
        "graph_all_positive": True,
    },
    "snn_na_padding": {
        "source": "tests/testthat/test-sparseSNN.R:71-79",
        "idx": [[2, 3], [1, None], [1, 2]], "include_self": True,
        "expect_w": [[2 / 3, 1.0], [2 / 3, None], [1.0, 2 / 3]],
    },
    "snn_out_of_range": {
        "source": "tests/testthat/test-sparseSNN.R:81-88",
        "cases": [[[1, 99]], [[1, 0]], [[1, -2]]], "error": "1-based row indices",
    },
    "snn_duplicates": {
        "source": "tests/testthat/test-sparseSNN.R:171-179",
        "idx": [[2, 2], [1, 1]], "include_self": True, "symmetrise": False,
        "expect_entries": [[1, 2, 1.0], [2, 1, 1.0]],
    },
    "snn_include_self_override": {
        "source": "tests/testthat/test-sparseSNN.R:218-238",
        "idx": [[1, 2], [2, 3], [3, 1]],
        "idx_noself": [[2, None], [3, None], [1, None]],
        "expect_with_self_1_2": 1 / 3, "expect_without_1_2": 0.0,
    },
    "snn_shared_awkward": {
        "source": "tests/testthat/test-sparseSNN-shared.R:26-37, 85-186",
        "idx": [[4, 2, None], [4, 1, None], [4, 5, 5], [4, 5, 6], [6, None, None],
                [5, None, None]],
        "expect_shared_pairs": 11,
        "expect_shared": [[5, 6, 1.0], [1, 3, 0.2], [3, 5, 0.25], [4, 5, 2 / 3]],
        "expect_knn_offdiag_pairs": 8,
        "expect_knn_selfloop": [4, 4, 1.0],
        "prune_values": [0.0, 0.2, 0.25, 0.5, 2 / 3],
        "prune_one_empty": True,
    },
    "js_standard_metric": {
        "source": "tests/testthat/test-sparseDist.R:100-115",
        # two point masses: d(a, b) = sqrt(log 2), the metric's upper bound
        "P": cols([1, 0], [0, 1]),
        "expect_d12": "sqrt(log(2))",
        "bound": "sqrt(log(2)) + 1e-12 for normalised columns",
    },
    "js_invalid_input": {
        "source": "tests/testthat/test-sparseDist.R:117-124",
        "cases": [{"M": [[-1.0, 2.0], [3.0, 4.0]], "error": "non-negative"},
                  {"M": [[None, 2.0], [3.0, 4.0]], "error": "finite"}],
    },
    "js_self_exact_zero": {
        "source": "tests/testthat/test-sparseKNN.R:271-277",
        "k": 2, "include_self": True,
        "expect": "dist[, 1] identical to 0 and idx[, 1] the column itself",
    },
    "distance_only_methods": {
        "source": "tests/testthat/test-sparseKNN.R:209-217",
        "methods": ["euclidean", "manhattan", "js"], "error": "only meaningful",
    },
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1)
    print("wrote", os.path.join(HERE, "fixtures.json"))
