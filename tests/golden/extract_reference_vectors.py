"""Writes reference_vectors.json: the golden inputs and expected outputs held inline in the
reference's gtest files, as data.

Run in the build container (where /root/reference exists):
    python tests/golden/extract_reference_vectors.py
The JSON it writes is committed; nothing at test time reads the reference tree.

Each entry records the reference file:line range it was read from.  Eigen's comma
initialiser (`M << a, b, ...`) fills row by row, so every matrix below is row-major.
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/DPMMSubClusters/test"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?"


def read(name):
    with open(os.path.join(REF, name)) as f:
        return f.read().split("\n")


def block(lines, start, end):
    """1-based inclusive line range as one string."""
    return "\n".join(lines[start - 1:end])


def nums(s):
    out = []
    for t in re.findall(NUM, s):
        if t.endswith("f"):
            out.append(float(np.float32(float(t[:-1]))))
        else:
            out.append(float(t))
    return out


def init_list(text, var, nth=0):
    """Values of the nth `var << ... ;` comma initialiser in text."""
    pat = re.compile(r"(?<![\w>])" + re.escape(var) + r"\s*<<(.*?);", re.S)
    found = pat.findall(text)
    return nums(found[nth])


def split_args(s):
    """Split a call's argument text at top-level commas."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
            continue
        depth += ch == "("
        depth -= ch == ")"
        cur += ch
    out.append(cur.strip())
    return out


def expects(text):
    """(expected, actual_expr, tol) of every EXPECT_NEAR / ASSERT_NEAR, in order."""
    out = []
    for m in re.finditer(r"(?:EXPECT|ASSERT)_NEAR\(", text):
        depth, j = 1, m.end()
        while depth:
            depth += text[j] == "("
            depth -= text[j] == ")"
            j += 1
        a = split_args(text[m.end():j - 1])
        out.append((nums(a[0])[0], a[1], nums(a[2])[0]))
    return out


def scalar_after(text, pattern):
    m = re.search(pattern + r"\s*(" + NUM + r")", text)
    return nums(m.group(1))[0]


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not mounted; the committed JSON is the fixture")
    g = {}

    # --- Gaussian label log-likelihood, weight 0.5 and weight 1 ---------------------------
    L = read("cudaKernel_gaussian_test.cpp")
    for key, (a, b) in {"gauss_ll_w05": (748, 787), "gauss_ll_w1": (789, 829)}.items():
        t = block(L, a, b)
        g[key] = {
            "source": f"test/cudaKernel_gaussian_test.cpp:{a}-{b}",
            "x_DxN": init_list(t, "x"), "D": 2, "N": 10,
            "mu": init_list(t, "mu"), "sigma": init_list(t, "sigma"),
            "inv_sigma": init_list(t, "invSigma"),
            "logdet_sigma": scalar_after(t, r"double logdetSigma\s*="),
            "weight": scalar_after(t, r"double weight\s*="),
            "expected": [e[0] for e in expects(t)], "tol": expects(t)[0][2],
        }

    # --- Multinomial label log-likelihood ----------------------------------------------------
    L = read("cudaKernel_multinomial_test.cpp")
    t = block(L, 18, 55)
    g["mult_ll"] = {
        "source": "test/cudaKernel_multinomial_test.cpp:18-55",
        "x_DxN": init_list(t, "x"), "D": 2, "N": 10,
        "logtheta": [float(v) for v in re.findall(r"alpha\.push_back\((" + NUM + r")\)", t)],
        "expected": [e[0] for e in expects(t)], "tol": expects(t)[0][2],
    }

    # --- NIW prior ------------------------------------------------------------------------------
    L = read("niw_test.cpp")
    t = block(L, 13, 35)
    g["niw_stats"] = {
        "source": "test/niw_test.cpp:13-35",
        "points_DxN": init_list(t, "points"), "D": 2, "N": 10,
        "expected_sum": [expects(t)[0][0], expects(t)[1][0]],
        "expected_S": [e[0] for e in expects(t)[2:6]], "tol": expects(t)[0][2],
    }
    t = block(L, 37, 58)
    g["niw_posterior"] = {
        "source": "test/niw_test.cpp:37-58",
        "k0": 1.0, "m0": init_list(t, "m"), "v0": 5.0, "psi0": init_list(t, "psi"),
        "N": 37196, "sum": init_list(t, "points_sum"), "S": init_list(t, "S"),
        "expected_psi": [e[0] for e in expects(t)], "tol": expects(t)[0][2],
    }
    for key, (a, b, D) in {"niw_logml_d3": (184, 203, 3), "niw_logml_d50": (205, 224, 50)}.items():
        t = block(L, a, b)
        post = re.search(r"posterior = std::make_shared<niw_hyperparams>\((" + NUM + r"), m, (" + NUM + r"), psi_posterior\)", t)
        n = re.search(r"make_shared<niw_sufficient_statistics>\((" + NUM + r"), points_sum", t)
        e = expects(t)[0]
        g[key] = {
            "source": f"test/niw_test.cpp:{a}-{b}", "D": D,
            "k0": 1.0, "v0": float(D), "psi0": "identity",
            "kn": nums(post.group(1))[0], "vn": nums(post.group(2))[0],
            "psin": init_list(t, "psi_posterior"), "N": int(nums(n.group(1))[0]),
            "expected": e[0], "tol": e[2],
        }

    # --- local_clusters_actions: stats + posteriors, split test, stats gather ---------------------
    L = read("local_clusters_actions_test.cpp")
    t = block(L, 15, 172)
    ex = expects(t)
    g["lca_update_stats"] = {
        "source": "test/local_clusters_actions_test.cpp:15-172",
        "points_DxN": init_list(t, "points"), "D": 2, "N": 10,
        "labels": [int(v) for v in re.search(r"labels = \{(.*?)\}", t).group(1).split(",")],
        "sub_labels": [int(v) for v in re.search(r"sub_labels = \{(.*?)\}", t).group(1).split(",")],
        "k0": 1.0, "m0": [0.0, 0.0], "v0": 5.0, "psi0": [1.0, 0.0, 0.0, 1.0],
        "expected": [[e[0], e[1]] for e in ex], "tol": ex[0][2],
    }
    t = block(L, 174, 279)
    sums = [init_list(t, "suff_statistics->points_sum", i) for i in range(3)]
    Ss = [init_list(t, "suff_statistics->S", i) for i in range(3)]
    Ns = [int(v) for v in re.findall(r"suff_statistics->N = (\d+);", t)]
    ms = [init_list(t, v) for v in ("m1", "m2", "m3")]
    psis = [init_list(t, v) for v in ("psi1", "psi2", "psi3")]
    hyp = re.findall(r"niw_hyperparams niwHyperparams\d\((" + NUM + r"), m\d, (" + NUM + r"), psi\d\)", t)
    g["lca_should_split"] = {
        "source": "test/local_clusters_actions_test.cpp:174-279",
        "alpha": 10.0, "uniform": 0.5, "k0": 1.0, "m0": [0.0, 0.0], "v0": 5.0, "psi0": [1.0, 0.0, 0.0, 1.0],
        "parts": [
            {"N": Ns[i], "sum": sums[i], "S": Ss[i], "kn": nums(hyp[i][0])[0], "vn": nums(hyp[i][1])[0],
             "mn": ms[i], "psin": psis[i]} for i in range(3)
        ],
        "expected_log_hr": 0.5208818912506104, "expected_split": 1,
        "note": "parts are cluster, left, right; log_HR value from the comment at line 277",
    }
    t = block(L, 281, 315)
    ex = expects(t)
    g["lca_stats_gather"] = {
        "source": "test/local_clusters_actions_test.cpp:281-315",
        "points_DxN": init_list(t, "group_pts"), "D": 2, "N": 10,
        "labels": [int(v) for v in re.search(r"labels = \{(.*?)\}", t).group(1).split(",")],
        "sub_labels": [int(v) for v in re.search(r"sub_labels = \{(.*?)\}", t).group(1).split(",")],
        "expected_Nl": 4, "expected_Nr": 6,
        "expected": [[e[0], e[1]] for e in ex], "tol": ex[0][2],
    }

    # --- multinomial prior ------------------------------------------------------------------------
    L = read("multinomial_prior_test.cpp")
    t = block(L, 14, 34)
    g["mult_stats"] = {
        "source": "test/multinomial_prior_test.cpp:14-34",
        "points_DxN": init_list(t, "points"), "D": 2, "N": 10,
        "expected_sum": [e[0] for e in expects(t)], "tol": expects(t)[0][2],
    }
    t = block(L, 55, 78)
    g["mult_logml"] = {
        "source": "test/multinomial_prior_test.cpp:55-78",
        "alpha0": init_list(t, "alpha_prior"), "alphan": init_list(t, "alpha_posterior"),
        "expected": expects(t)[0][0], "tol": expects(t)[0][2],
    }
    t = block(L, 36, 53)
    g["mult_draw"] = {
        "source": "test/multinomial_prior_test.cpp:36-53",
        "alpha": init_list(t, "alpha"), "expected_logtheta": [e[0] for e in expects(t)], "tol": expects(t)[0][2],
    }

    # --- shared_actions: burn-out history of sample_cluster_params ---------------------------------
    L = read("shared_actions_test.cpp")
    t = block(L, 78, 418)
    # the left/right states fed to sample_cluster_params: iterations 1-2 (initial), 3 and 4
    init_l = {"N": 6, "kn": 7.0, "vn": 11.0, "psin": init_list(t, "psi2")}
    init_r = {"N": 4, "kn": 5.0, "vn": 9.0, "psin": init_list(t, "psi3")}
    it3 = block(L, 352, 380)
    it4 = block(L, 381, 410)
    st3 = [init_list(it3, "nh->psi", i) for i in range(2)]
    st4 = [init_list(it4, "nh->psi", i) for i in range(2)]
    g["sa_hist"] = {
        "source": "test/shared_actions_test.cpp:78-418",
        "k0": 1.0, "v0": 5.0, "psi0": [1.0, 0.0, 0.0, 1.0], "burnout": 15,
        "states": [
            [init_l, init_r],
            [{"N": 7, "kn": 8.0, "vn": 12.0, "psin": st3[0]}, {"N": 3, "kn": 4.0, "vn": 8.0, "psin": st3[1]}],
            [{"N": 6, "kn": 7.0, "vn": 11.0, "psin": st4[0]}, {"N": 4, "kn": 5.0, "vn": 9.0, "psin": st4[1]}],
        ],
        "expected_hist": [e[0] for e in expects(block(L, 414, 417))],
        "tol": 1e-4,
        "note": "hist[0..1] from the initial state, hist[2] after the iteration-3 state, hist[3] after iteration 4",
    }

    # --- sampler statistics: categorical draw counts (distribution only) ---------------------------
    L = read("cudaKernel_gaussian_test.cpp")
    t = block(L, 18, 70)
    g["cat_draw_counts"] = {
        "source": "test/cudaKernel_gaussian_test.cpp:18-70",
        "log_likelihood": [0.4, 0.6], "n": 100000,
        "reference_counts": [45260, 54740],
        "note": "exact counts are cuRAND-XORWOW specific; only the distribution transfers",
    }

    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", OUT, "entries:", len(g))


if __name__ == "__main__":
    main()
