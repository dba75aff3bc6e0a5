"""Writes the golden fixtures transcribed from the reference's own C++ tests.

Each fixture is data: the input literals of a reference test and the expected
result that test asserts, as the test states it.  Nothing here runs or imports
the reference; the literals were read from its test sources, cited per fixture.

Run:  python tests/golden/make_fixtures.py   (rewrites tests/golden/*.json)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# --------------------------------------------------------------------------- split
# native-sql-engine/cpp/src/tests/shuffle_split_test.cc:97-208, SplitterTest fixture
SPLIT_SCHEMA = [
    ("f_na", "null"), ("f_int8_a", "int8"), ("f_int8_b", "int8"), ("f_int32", "int32"),
    ("f_uint64", "uint64"), ("f_double", "double"), ("f_bool", "bool"), ("f_string", "utf8"),
    ("f_nullable_string", "utf8"), ("f_decimal128", "decimal128(10,2)"),
]
INPUT_DATA_1 = [  # shuffle_split_test.cc:186-196
    [None] * 10,
    [1, 2, 3, None, 4, None, 5, 6, None, 7],
    [1, -1, None, None, -2, 2, None, None, 3, -3],
    [1, 2, 3, 4, None, 5, 6, 7, 8, None],
    [None] * 10,
    [-0.1234567, None, 0.1234567, None, -0.142857, None, 0.142857, 0.285714, 0.428617, None],
    [None, True, False, None, True, True, False, True, None, None],
    ["alice0", "bob1", "alice2", "bob3", "Alice4", "Bob5", "AlicE6", "boB7", "ALICE8", "BOB9"],
    ["alice", "bob", None, None, "Alice", "Bob", None, "alicE", None, "boB"],
    ["-1.01", "2.01", "-3.01", None, "0.11", "3.14", "2.27", None, "-3.14", None],
]
INPUT_DATA_2 = [  # shuffle_split_test.cc:198-208
    [None, None],
    [None, None],
    [1, -1],
    [100, None],
    [1, 1],
    [0.142857, -0.142857],
    [True, False],
    ["bob", "alice" * 40],
    [None, None],
    [None, None],
]

SPLIT_FIXTURES = {
    "schema": SPLIT_SCHEMA,
    "batches": {"input_batch_1": INPUT_DATA_1, "input_batch_2": INPUT_DATA_2},
    "tests": [
        {   # TestSingleSplitter, shuffle_split_test.cc:210-253
            "name": "TestSingleSplitter", "short_name": "rr", "num_partitions": 1,
            "buffer_size": 10,
            "split": ["input_batch_1", "input_batch_2", "input_batch_1"],
            # one partition; its record batches equal the three inputs
            "expected": [[["input_batch_1", list(range(10))], ["input_batch_2", [0, 1]],
                          ["input_batch_1", list(range(10))]]],
        },
        {   # TestRoundRobinSplitter, :255-319
            "name": "TestRoundRobinSplitter", "short_name": "rr", "num_partitions": 2,
            "buffer_size": 4,
            "split": ["input_batch_1", "input_batch_2", "input_batch_1"],
            "expected": [
                [["input_batch_1", [0, 2, 4, 6, 8]], ["input_batch_2", [0]],
                 ["input_batch_1", [0, 2, 4, 6, 8]]],
                [["input_batch_1", [1, 3, 5, 7, 9]], ["input_batch_2", [1]],
                 ["input_batch_1", [1, 3, 5, 7, 9]]],
            ],
        },
        {   # TestFallbackRangeSplitter, :390-472 (pid column prepended)
            "name": "TestFallbackRangeSplitter", "short_name": "range", "num_partitions": 2,
            "buffer_size": 4,
            "split": ["input_batch_1", "input_batch_2", "input_batch_1"],
            "pids": {"input_batch_1": [0, 1, 0, 1, 0, 1, 0, 1, 0, 1], "input_batch_2": [0, 1]},
            "expected": [
                [["input_batch_1", [0, 2, 4, 6, 8]], ["input_batch_2", [0]],
                 ["input_batch_1", [0, 2, 4, 6, 8]]],
                [["input_batch_1", [1, 3, 5, 7, 9]], ["input_batch_2", [1]],
                 ["input_batch_1", [1, 3, 5, 7, 9]]],
            ],
        },
    ],
}


# --------------------------------------------------------------------------- aggregate
# native-sql-engine/cpp/src/tests/arrow_compute_test_aggregate.cc.  Doubles that
# the tests print rounded (6 significant digits) carry "tolerance".
NAN = "NaN"
AGG_FIXTURES = {"tests": [
    {   # MinMaxNaNTest, :222-283 (global min/max with NaN check)
        "name": "MinMaxNaNTest",
        "schema": [["f0", "double"]],
        "actions": [["action_min_true", [0]], ["action_max_true", [0]]],
        "batches": [
            [[1, 2, 3, 4, 5, NAN, 6, 7, 8, 9, 10, 16, 19, 42, 78, 12, 5, NAN, 11, 19]],
            [[16, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 18, 19, 12, 13]],
        ],
        "expected": [[1], [NAN]],
    },
    {   # GroupByMinMaxNaNTest, :284-364
        "name": "GroupByMinMaxNaNTest",
        "schema": [["f0", "int64"], ["f1", "double"]],
        "actions": [["action_groupby", [0]], ["action_min_true", [1]], ["action_max_true", [1]]],
        "batches": [
            [[1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [1, 2, 3, 4, 5, NAN, 6, 7, 8, 9, 10, 16, 19, 42, 78, 12, 5, NAN, 11, 19]],
            [[6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [16, 7, 8, 9, 10, NAN, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 18, 19, 12, 13]],
        ],
        "expected": [[1, 2, 3, 4, 5, None, 6, 7, 8, 9, 10],
                     [1, 2, 3, 4, 5, NAN, 6, 7, 8, 9, 10],
                     [19, 9, 12, 78, NAN, NAN, 16, 7, 8, 9, NAN]],
    },
    {   # AggregateAllNullTest, :365-421
        "name": "AggregateAllNullTest",
        "schema": [["f0", "int32"]],
        "actions": [["action_sum", [0]]],
        "batches": [[[None] * 12], [[None] * 12]],
        "expected": [[None]],
    },
    {   # GroupByAggregateTest, :422-547: the in-scope actions; the test's result
        # expression multiplies sum by count ("sum_count_multiply")
        "name": "GroupByAggregateTest",
        "schema": [["f0", "int64"], ["f1", "uint32"], ["f2", "double"], ["f3", "double"], ["f4", "double"]],
        "actions": [["action_groupby", [0]], ["action_sum_count", [1]], ["action_min", [1]],
                    ["action_max", [1]], ["action_avgByCount", [2, 0]]],
        "batches": [
            [[1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [1, 2, 3, 4, 5, 5, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1]],
            [[6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [7, 8, 4, 5, 6, 1, 34, 54, 65, 66, 78, 12, 32, 24, 32, 45, 12, 24, 35, 46],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1]],
        ],
        "expected_named": {
            "unique": [1, 2, 3, 4, 5, None, 6, 7, 8, 9, 10],
            "sum_count_multiply": [25, 18, 12, 64, 125, 5, 150, 63, 32, 144, 360],
            "min": [1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10],
            "avg": [16.4, 6.5, 5, 5.875, 5.48, None, 6.1, 6.61905, 3.0625, 2.63889, 2.06667],
        },
        "tolerance": 1e-5,
    },
    {   # GroupByTwoAggregateTest, :911-1040: GROUP BY (f0 int64, f5 utf8).  Its
        # strings are distinct non-empty 2-byte values, so the UnsafeRow byte key
        # of (f0, f5) is equal exactly when (f0, code(f5)) is: f5 is given here as
        # the int32 code BJ 0, SH 1, HZ 2, BH 3, NY 4, TK 5, PH 6.  The
        # stddev_samp action is left out.
        "name": "GroupByTwoAggregateTest",
        "schema": [["f0", "int64"], ["f1", "uint32"], ["f2", "double"], ["f3", "double"], ["f4", "double"],
                   ["f5_code", "int32"]],
        "actions": [["action_groupby", [0]], ["action_groupby", [5]], ["action_sum_count", [1]],
                    ["action_min", [1]], ["action_max", [1]], ["action_avgByCount", [2, 0]]],
        "batches": [
            [[1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [1, 2, 3, 4, 5, 5, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [0, 1, 2, 3, 4, 1, 3, 0, 1, 1, 0, 0, 0, 3, 3, 2, 4, 4, 4, 4]],
            [[6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [7, 8, 4, 5, 6, 1, 34, 54, 65, 66, 78, 12, 32, 24, 32, 45, 12, 24, 35, 46],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [0, 1, 5, 1, 6, 6, 1, 0, 1, 1, 0, 0, 0, 1, 1, 5, 6, 6, 6, 6]],
        ],
        "expected": [[1, 2, 3, 4, 5, None, 6, 7, 8, 9, 10],                                 # :1020-1031
                     [0, 1, 2, 3, 4, 1, 0, 1, 5, 1, 6],
                     [5, 6, 6, 16, 25, 5, 30, 21, 16, 36, 60],
                     [5, 3, 2, 4, 5, 1, 5, 3, 2, 4, 6],
                     [1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10],
                     [1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10],
                     [16.4, 6.5, 5, 5.875, 5.48, None, 6.1, 6.61905, 3.0625, 2.63889, 2.06667]],
        "tolerance": 1e-5,
    },
    {   # GroupByTwoAggregateTest, :911-1040, with its utf8 key column as written
        # (the stddev_samp action left out)
        "name": "GroupByTwoAggregateTestUtf8",
        "schema": [["f0", "int64"], ["f1", "uint32"], ["f2", "double"], ["f3", "double"], ["f4", "double"],
                   ["f5", "utf8"]],
        "actions": [["action_groupby", [0]], ["action_groupby", [5]], ["action_sum_count", [1]],
                    ["action_min", [1]], ["action_max", [1]], ["action_avgByCount", [2, 0]]],
        "batches": [
            [[1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [1, 2, 3, 4, 5, 5, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             ["BJ", "SH", "HZ", "BH", "NY", "SH", "BH", "BJ", "SH", "SH", "BJ", "BJ", "BJ", "BH", "BH", "HZ",
              "NY", "NY", "NY", "NY"]],
            [[6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [7, 8, 4, 5, 6, 1, 34, 54, 65, 66, 78, 12, 32, 24, 32, 45, 12, 24, 35, 46],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             ["BJ", "SH", "TK", "SH", "PH", "PH", "SH", "BJ", "SH", "SH", "BJ", "BJ", "BJ", "SH", "SH", "TK",
              "PH", "PH", "PH", "PH"]],
        ],
        "expected": [[1, 2, 3, 4, 5, None, 6, 7, 8, 9, 10],                                 # :1020-1031
                     ["BJ", "SH", "HZ", "BH", "NY", "SH", "BJ", "SH", "TK", "SH", "PH"],
                     [5, 6, 6, 16, 25, 5, 30, 21, 16, 36, 60],
                     [5, 3, 2, 4, 5, 1, 5, 3, 2, 4, 6],
                     [1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10],
                     [1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10],
                     [16.4, 6.5, 5, 5.875, 5.48, None, 6.1, 6.61905, 3.0625, 2.63889, 2.06667]],
        "tolerance": 1e-5,
    },
    {   # AggregateTest, :30-221 (global aggregate): the in-scope actions.  Left
        # out: action_stddev_samp_final, action_min/max over utf8 and
        # action_first_partial/final, which are not on the built path.
        "name": "AggregateTest",
        "schema": [["f0", "int32"], ["f1", "int32"], ["f2", "double"], ["f3", "double"], ["f4", "double"],
                   ["f5", "utf8"], ["f6", "utf8"], ["f7", "bool"]],
        "actions": [["action_sum", [0]], ["action_count", [0]], ["action_sum_count", [0]],
                    ["action_avgByCount", [2, 1]], ["action_min", [0]], ["action_max", [0]],
                    ["action_countLiteral_1", []]],
        "batches": [
            [[1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],           # :160-178
             [1, 2, 3, 4, 5, None, 4, 1, 2, 2, 1, 1, 1, 4, 4, 3, 5, 5, 5, 5],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             ["BJK", "SH", "SZ", "NY", "WH", "WH", "AU", "BJ", "SH", "DL", "cD", "CD", "Bj", "LA", "HZ", "LA",
              "WH", "NY", "WH", "WH"],
             ["ZZ", "SH", "SZ", "NY", "WH", "WH", "AU", "BJ", "SH", "DL", "cD", "CD", "Bj", "LA", "HZ", "LA",
              "WH", "NY", "WH", "WH"],
             [True, False, True, False, False, True, False, True, False, False, True, False, True, False,
              False, True, False, True, False, False]],
            [[6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],          # :180-199
             [6, 7, 8, 9, 10, 10, 9, 6, 7, 7, 6, 6, 6, 9, 9, 8, 10, 10, 10, 10],
             [7, 8, 4, 5, 6, 1, 34, 54, 65, 66, 78, 12, 32, 24, 32, 45, 12, 24, 35, 46],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             [2, 4, 5, 7, 8, 2, 45, 32, 23, 12, 14, 16, 18, 19, 23, 25, 57, 59, 12, 1],
             ["BJ", "SHL", "SZ", "NY", "WH", "WH", "au", "BJ", "SH", "DL", "cD", "CD", "Bj", "LA", "HZ", "LA",
              "WH", "NY", "WHZ", "wH"],
             ["BJ", "SHL", "SZ", "NY", "WH", "WH", "au", "BJ", "SH", "DL", "cD", "CD", "Bj", "LA", "HZ", "LA",
              "WH", "NY", "WHZ", "wH"],
             [True, False, True, False, False, True, False, True, False, False, True, False, True, False,
              False, True, False, True, False, False]],
        ],
        # :203-207: sum, count, sum_count (sum, count), avgByCount, min, max, count_all
        "expected": [[221], [39], [221], [39], [4.40724], [1], [10], [40]],
        "tolerance": 1e-5,
    },
    {   # GroupByCountAll, :755-832: GROUP BY a utf8 key, countLiteral_1
        "name": "GroupByCountAll",
        "schema": [["f0", "utf8"]],
        "actions": [["action_groupby", [0]], ["action_countLiteral_1", []]],
        "batches": [
            [["BJ", "SH", "SZ", "HZ", "WH", "WH", "HZ", "BJ", "SH", "SH", "BJ", "BJ", "BJ", "HZ", "HZ", "SZ",
              "WH", "WH", "WH", "WH"]],
            [["CD", "DL", "NY", "LA", "AU", "AU", "LA", "CD", "DL", "DL", "CD", "CD", "CD", "LA", "LA", "NY",
              "AU", "AU", "AU", "AU"]],
            [["BJ", "SH", "SZ", "NY", "WH", "WH", "AU", "BJ", "SH", "DL", "CD", "CD", "BJ", "LA", "HZ", "LA",
              "WH", "NY", "WH", "WH"]],
        ],
        "expected": [["BJ", "SH", "SZ", "HZ", "WH", "CD", "DL", "NY", "LA", "AU"],         # :823-825
                     [8, 5, 3, 5, 11, 7, 4, 4, 6, 7]],
    },
    {   # GroupByCountOnMutipleCols, :833-910: GROUP BY a utf8 key, count over
        # two utf8 columns (rows where both are non-null)
        "name": "GroupByCountOnMutipleCols",
        "schema": [["f0", "utf8"], ["f1", "utf8"], ["f2", "utf8"]],
        "actions": [["action_groupby", [0]], ["action_count", [1, 2]]],
        "batches": [
            [["a", "a", "a", "x", "x"], ["b", "b", "b", "y", "q"], [None, "c", "d", "z", None]],
            [["b", "a", "b", "a", "x"], ["b", "b", "b", None, "q"], ["c", None, "d", "z", None]],
        ],
        "expected": [["a", "x", "b"], [2, 1, 2]],                                          # :901
    },
]}


# --------------------------------------------------------------------------- join
# native-sql-engine/cpp/src/tests/arrow_compute_test_join_wocg.cc.  Results list
# the output columns per probe batch; the result schema holds probe ("right")
# columns, as each test's "result" node does.
_BUILD = [[[10, 3, 1, 2, 3, 1], [10, 3, 1, 2, 13, 11], [10, 3, 1, 2, 13, 11]],
          [[6, 12, 5, 8, 6, 10], [6, 12, 5, 8, 16, 110], [6, 12, 5, 8, 16, 110]]]
_PROBE_STR = [[[1, 3, 4, 5, 6], ["BJ", "TY", "NY", "SH", "HZ"]],
              [[7, 8, 9, 10, 11, 12], ["SH", "NY", "BJ", "IT", "BR", "TL"]]]
JOIN_FIXTURES = {"tests": [
    {   # JoinWOCGTestAntiJoinType2, :1592-1712
        "name": "AntiJoinType2", "join_type": "anti", "null_aware": False,
        "build_schema": [["table0_f0", "uint32"], ["table0_f1", "uint32"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "uint32"], ["table1_f1", "uint32"]],
        "build_key": 0, "probe_key": 0, "result": [[1, 0], [1, 1]],
        "build": _BUILD,
        "probe": [[[1, 2, 3, 4, 5, 6], [1, 2, 3, 4, 5, 6]], [[7, 8, 9, 10, 11, 12], [7, 8, 9, 10, 11, 12]]],
        "expected": [[[4], [4]], [[7, 9, 11], [7, 9, 11]]],
    },
    {   # JoinWOCGTestAntiJoinSingleKey (null-aware anti), :1714-1818
        "name": "AntiJoinSingleKey", "join_type": "anti", "null_aware": True,
        "build_schema": [["table0_f0", "int32"], ["table0_f1", "int32"]],
        "probe_schema": [["table1_f0", "int32"], ["table1_f1", "int32"]],
        "build_key": 0, "probe_key": 0, "result": [[1, 0], [1, 1]],
        "build": [[[2, 2, 3, 4, None, None, 6], [3, 3, 2, 1, None, 5, None]]],
        "probe": [[[1, 1, 2, 2, 3, None, None, 6], [2, 2, 1, 1, 3, None, 5, None]]],
        "expected": [[[], []]],
    },
    {   # JoinWOCGTestSemiJoinType3 (boolean key), :2056-2154
        "name": "SemiJoinType3", "join_type": "semi", "null_aware": False,
        "build_schema": [["table0_f0", "bool"]],
        "probe_schema": [["table1_f0", "int32"], ["table1_f1", "bool"]],
        "build_key": 0, "probe_key": 1, "result": [[1, 0], [1, 1]],
        "build": [[[True]]],
        "probe": [[[2], [True]]],
        "expected": [[[2], [True]]],
    },
    {   # JoinWOCGTestSemiJoinType2WithUInt64, :2283-2403
        "name": "SemiJoinType2WithUInt64", "join_type": "semi", "null_aware": False,
        "build_schema": [["table0_f0", "uint64"], ["table0_f1", "uint32"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "uint64"], ["table1_f1", "utf8"]],
        "build_key": 0, "probe_key": 0, "result": [[1, 0], [1, 1]],
        "build": _BUILD, "probe": _PROBE_STR,
        "expected": [[[1, 3, 5, 6], ["BJ", "TY", "SH", "HZ"]], [[8, 10, 12], ["NY", "IT", "TL"]]],
    },
    {   # JoinWOCGTestSemiJoinType2, :1934-2054 (uint32 key)
        "name": "SemiJoinType2", "join_type": "semi", "null_aware": False,
        "build_schema": [["table0_f0", "uint32"], ["table0_f1", "uint32"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "uint32"], ["table1_f1", "utf8"]],
        "build_key": 0, "probe_key": 0, "result": [[1, 0], [1, 1]],
        "build": _BUILD, "probe": _PROBE_STR,
        "expected": [[[1, 3, 5, 6], ["BJ", "TY", "SH", "HZ"]], [[8, 10, 12], ["NY", "IT", "TL"]]],
    },
    {   # JoinWOCGTestAntiJoinMulKeys, :1820-1932: conditionedProbeArraysAnti_true
        # (null-aware) over two int32 keys (UnsafeRow keys, hash config 1)
        "name": "AntiJoinMulKeys", "join_type": "anti", "null_aware": True,
        "build_schema": [["table0_f0", "int32"], ["table0_f1", "int32"], ["table0_f2", "int32"]],
        "probe_schema": [["table1_f0", "int32"], ["table1_f1", "int32"], ["table1_f2", "int32"]],
        "build_key": [0, 1], "probe_key": [0, 1], "result": [[1, 0], [1, 1]],
        "build": [[[1, None, None, 1, 1], [1, None, 3, None, 2], [2, 2, 2, 2, None]]],
        "probe": [[[1, 2, None, 1, None, None, 1, None], [1, 1, 1, 3, None, 3, None, 2], [2, 3, 6, 1, 6, 2, 9, 2]]],
        "expected": [[[2, None, 1, None, None, 1, None], [1, 1, 3, None, 3, None, 2]]],
    },
    {   # JoinWOCGTestInnerJoinNaN, :1349-1462: two double keys, each wrapped in
        # "normalize" (normalize_nan_zero), UnsafeRow keys; result: both build
        # columns, then both probe columns
        "name": "InnerJoinNaN", "join_type": "inner", "null_aware": False,
        "build_schema": [["table0_f0", "double"], ["table0_f1", "double"]],
        "probe_schema": [["table1_f0", "double"], ["table1_f1", "double"]],
        "build_key": [0, 1], "probe_key": [0, 1], "normalize": [True, True],
        "result": [[0, 0], [0, 1], [1, 0], [1, 1]],
        "build": [[[NAN, 0.0, -0.0], [NAN, 0.0, -0.0]]],
        "probe": [[[NAN, 0.0, -0.0], [NAN, 0.0, -0.0]]],
        "expected": [[[NAN, 0.0, -0.0, 0.0, -0.0], [NAN, 0.0, -0.0, 0.0, -0.0],
                      [NAN, 0.0, 0.0, -0.0, -0.0], [NAN, 0.0, 0.0, -0.0, -0.0]]],
    },
    {   # JoinWOCGTestStringInnerJoin, :95-225: a utf8 key; the probe key is the
        # projection upper(table1_f0), given here as the probe's third column
        # (the value the reference's key projection computes before probing)
        "name": "StringInnerJoin", "join_type": "inner", "null_aware": False,
        "build_schema": [["table0_f0", "utf8"], ["table0_f1", "utf8"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "utf8"], ["table1_f1", "uint32"], ["upper_table1_f0", "utf8"]],
        "build_key": 0, "probe_key": 2, "result": [[0, 0], [0, 1], [0, 2], [1, 0], [1, 1]],
        "build": [[["BJ", "SH", "HZ", "BH", "NY", "SH"], ["A", "A", "C", "D", "C", "D"], [10, 3, 1, 2, 13, 11]],
                  [["TK", "SH", "PH", "NJ", "NB", "SZ"], ["F", "F", "A", "B", "D", "C"], [6, 12, 5, 8, 16, 110]]],
        "probe": [[["sh", "sz", "bj", None, "ny", "hz"], [1, 2, 3, 4, 5, 6],
                   ["SH", "SZ", "BJ", None, "NY", "HZ"]],
                  [["ph", None, "jh", "kk", "nj", "sz"], [7, 8, 9, 10, None, 12],
                   ["PH", None, "JH", "KK", "NJ", "SZ"]]],
        "expected": [[["SH", "SH", "SH", "SZ", "BJ", "NY", "HZ"], ["A", "D", "F", "C", "A", "C", "C"],
                      [3, 11, 12, 110, 10, 13, 1], ["sh", "sh", "sh", "sz", "bj", "ny", "hz"],
                      [1, 1, 1, 2, 3, 5, 6]],
                     [["PH", "NJ", "SZ"], ["A", "B", "C"], [5, 8, 110], ["ph", "nj", "sz"], [7, None, 12]]],
    },
    # JoinWOCGTestProjectKeyInnerJoin (:41-94) only builds the probe's code
    # generator: it holds no input rows and no expected result to transcribe.
    {   # JoinWOCGTestInnerJoinType2WithUInt16 (repeated build keys), :2405-2532
        "name": "InnerJoinType2WithUInt16", "join_type": "inner", "null_aware": False,
        "build_schema": [["table0_f0", "uint16"], ["table0_f1", "uint32"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "uint16"], ["table1_f1", "utf8"]],
        "build_key": 0, "probe_key": 0, "result": [[1, 0], [1, 1]],
        "build": _BUILD, "probe": _PROBE_STR,
        "expected": [[[1, 1, 3, 3, 5, 6, 6], ["BJ", "BJ", "TY", "TY", "SH", "HZ", "HZ"]],
                     [[8, 10, 10, 12], ["NY", "IT", "IT", "TL"]]],
    },
]}


# --------------------------------------------------------------------------- whole stage
# native-sql-engine/cpp/src/tests/arrow_compute_test_wscg.cc:4838-5000,
# WSCGTestInnerJoinWithGroupbyAggregate: table0 is the build side (key f0),
# table1 probes it (key f0); the join condition is add(table0_f1, table1_f1) >
# table0_f2; the aggregate groups by table0_f0 with sum_count(table1_f1) and
# min(table1_f1).
WSCG_FIXTURES = {"tests": [
    {
        "name": "WSCGTestInnerJoinWithGroupbyAggregate",
        "build_schema": [["table0_f0", "uint32"], ["table0_f1", "uint32"], ["table0_f2", "uint32"]],
        "probe_schema": [["table1_f0", "uint32"], ["table1_f1", "uint32"]],
        "build_key": 0, "probe_key": 0,
        "condition": "table0_f1 + table1_f1 > table0_f2",
        "aggregate_inputs": [[0, 0], [1, 1]],  # (side, column): table0_f0, table1_f1
        "actions": [["action_groupby", [0]], ["action_sum_count", [1]], ["action_min", [1]]],
        "build": [[[10, 3, 1, 2, 3, 1], [10, 3, 1, 2, 13, 11], [10, 3, 1, 2, 13, 11]],   # :4940-4946
                  [[6, 12, 5, 8, 6, 10], [6, 12, 5, 8, 16, 110], [6, 12, 5, 8, 16, 110]]],
        "probe": [[[1, 2, 3, 4, 5, 6], [1, 2, 3, 4, 5, 6]],                            # :4948-4954
                  [[7, 8, 9, 10, 11, 12], [7, 8, 9, 10, 11, 12]]],
        "result_schema": [["table0_f0_unique", "uint32"], ["table1_f1_sum", "double"],
                          ["table1_f1_count", "int64"], ["table1_f1_min", "uint32"]],
        "expected": [[1, 2, 3, 5, 6, 8, 10, 12], [2, 2, 6, 5, 12, 8, 20, 12],           # :4958-4962
                     [2, 1, 2, 1, 2, 1, 2, 1], [1, 2, 3, 5, 6, 8, 10, 12]],
    },
]}


def main():
    with open(os.path.join(HERE, "split_fixtures.json"), "w") as f:
        json.dump(SPLIT_FIXTURES, f, indent=1)
    with open(os.path.join(HERE, "agg_fixtures.json"), "w") as f:
        json.dump(AGG_FIXTURES, f, indent=1)
    with open(os.path.join(HERE, "join_fixtures.json"), "w") as f:
        json.dump(JOIN_FIXTURES, f, indent=1)
    with open(os.path.join(HERE, "wscg_fixtures.json"), "w") as f:
        json.dump(WSCG_FIXTURES, f, indent=1)
    print("wrote split_fixtures.json, agg_fixtures.json, join_fixtures.json, wscg_fixtures.json")


if __name__ == "__main__":
    main()
