"""Shows that the checks pinning the oracle are not vacuous.

Each case below plants one plausible mistake in a copy of oracle/oracle.cpp
(a dropped term, a wrong label, an inverted test, an off-by-one), builds that
copy on the CPU, and runs the pinning tests of test_oracle_mce.py or
test_oracle_gm.py against it.  At least one of them must fail.  The mistakes
are planted only in the restatements (orc_mce, orc_gm, orc_two_hop and their
helpers), never in the independent checkers they are compared with.  The cost model's CPU
reference (oracle/mlp_ref.py) gets the same treatment against test_mlp_ref.py.
"""
import importlib.util
import json
import os
import subprocess

import pytest

import test_mlp_ref as TL
import test_oracle_features as TF
import test_oracle_gm as TG
import test_oracle_mce as TM

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle", "oracle.cpp")
CHECKERS_MARK = "// Independent checkers (brute force)"

# (name, suite, original text, planted text)
MUTANTS = [
    # app_BK/BK.h:78-91: a node is a maximal clique only when P and X are empty
    ("mce_clique_test_ignores_x", "mce",
     "if (e.l == LP || e.l == LX) { anyPX = true; break; }",
     "if (e.l == LP) { anyPX = true; break; }"),
    # app_BK/BK.h:282-321: ties go to the last maximum
    ("mce_pivot_first_maximum", "mce",
     "if (cnt >= maxc) { maxc = cnt; pivot = u; }",
     "if (cnt > maxc || nPX == 1) { maxc = cnt; pivot = u; }"),
    # app_BK/BK.h:282-321: the pivot ranges over P and X
    ("mce_pivot_from_p_only", "mce",
     "        if (e.l == LR) continue;\n        nPX++;",
     "        if (e.l != LP) continue;\n        nPX++;"),
    # app_BK/BK.h:445-464: Q = P minus N(pivot)
    ("mce_q_inverted", "mce",
     "if (e.l == LP && !member(g.cols, plo, phi, e.v)) e.l = LQ;",
     "if (e.l == LP && member(g.cols, plo, phi, e.v)) e.l = LQ;"),
    # app_BK/BK.h:236-280: earlier Q siblings move to X in the child
    ("mce_earlier_siblings_stay_in_p", "mce",
     "if (l == LQ) l = (e.v < q) ? LX : LP;",
     "if (l == LQ) l = LP;"),
    # app_BK/BK.h:323-355: neighbours below the root start in X
    ("mce_root_without_x", "mce",
     "root.push_back({u, u < v ? LX : LP});",
     "root.push_back({u, LP});"),
    # histogram bin = clique size
    ("mce_histogram_off_by_one", "mce",
     "uint64_t b = std::min<uint64_t>(nR, acc.hist.size() - 1);",
     "uint64_t b = std::min<uint64_t>(nR - 1, acc.hist.size() - 1);"),
    # app_BK/BK.h:335-336: degree-0 roots are skipped
    ("mce_isolated_roots_counted", "mce",
     "if (hi > lo) {  // deg-0 roots",
     "if (true) {  // deg-0 roots"),
    # app_gmatch/gmatch.h:559-588: NON_EQUAL conditions
    ("gm_non_equal_ignored", "gm",
     "else { if (m == w) return false; }",
     "else { }"),
    # app_gmatch/gmatch.h:559-588: LARGER is strict
    ("gm_larger_not_strict", "gm",
     "else if (op == C_LARGER) { if (m >= w) return false; }",
     "else if (op == C_LARGER) { if (m > w) return false; }"),
    # app_gmatch/gmatch.h:589-601: membership in every other back neighbour
    ("gm_membership_skipped", "gm",
     "            const uint32_t bm = c.M[c.p->rev[bns[j]]];\n"
     "            ok = member(c.g.cols, c.g.rp[bm], c.g.rp[bm + 1], w);",
     "            const uint32_t bm = c.M[c.p->rev[bns[j]]];\n"
     "            ok = bm == bm;"),
    # app_gmatch/gmatch.h:273-509: shared positions filter by the after set
    ("gm_shared_after_membership_skipped", "gm",
     "                const uint32_t bm = c.M[p->rev[p->after_bn[id][j]]];\n"
     "                ok = member(c.g.cols, c.g.rp[bm], c.g.rp[bm + 1], w);",
     "                const uint32_t bm = c.M[p->rev[p->after_bn[id][j]]];\n"
     "                ok = bm == bm;"),
    ("gm_shared_after_conditions_skipped", "gm",
     "if (!conds_ok(c, p->after_cond_count[u], p->after_cond[u], w)) continue;",
     "(void)u;"),
    ("gm_shared_pre_conditions_skipped", "gm",
     "p->pre_cond_count[u], p->pre_cond[u], pre);",
     "0, p->pre_cond[u], pre);"),
    # the last position adds every candidate, not one per prefix
    ("gm_last_level_counts_prefixes", "gm",
     "if (id + 1 == p->n) { c.count += cands.size(); return; }",
     "if (id + 1 == p->n) { c.count += !cands.empty(); return; }"),
    # the byte model: c(A,u) = 4 min(|A| + d(u), |A| (1 + bitlen d(u)))
    ("bytes_merge_cost_off_by_one", "mce",
     "uint64_t merge = a + d;",
     "uint64_t merge = a + d + 1;"),
    # BK node: 5s for reading the node
    ("mce_bytes_node_read", "mce",
     "acc.bytes += 5 * s;",
     "acc.bytes += 4 * s;"),
    # BK node: c(P, u*) for Q = P minus N(u*)
    ("mce_bytes_pivot_term_dropped", "mce",
     "acc.bytes += cost_isect(pl.size(), phi - plo);",
     "(void)phi;"),
    # BK node: each child scans P u X against N(q)
    ("mce_bytes_child_scan_over_p", "mce",
     "acc.bytes += cost_isect(nPX, qhi - qlo) + 5 * child.size();",
     "acc.bytes += cost_isect(pl.size(), qhi - qlo) + 5 * child.size();"),
    # GM extension: membership tests per candidate passing the conditions
    ("gm_bytes_membership_dropped", "gm",
     "c.bytes += bcost;",
     "(void)bcost;"),
    # per-root outputs are each root's own subtree, not a running total
    ("mce_per_root_nodes_cumulative", "mce",
     "if (per_root_nodes) per_root_nodes[i] = acc.nodes - n0;",
     "if (per_root_nodes) per_root_nodes[i] = acc.nodes;"),
    ("gm_per_root_counts_cumulative", "gm",
     "if (per_root_count) per_root_count[i] = c.count - c0;",
     "if (per_root_count) per_root_count[i] = c.count;"),
    # GM extension: an EXPAND list stores whole partial matches
    ("gm_bytes_expand_priced_as_prefix", "gm",
     "return 4 * (uint64_t)(id + 1) * nout;",
     "return 4 * (id + nout);"),
    # 2-hop ball estimate: rank = leading zeros + 1
    ("twohop_rank_off_by_one", "feat",
     "std::min<uint32_t>((uint32_t)__builtin_clzll(rest) + 1, 32u)",
     "std::min<uint32_t>((uint32_t)__builtin_clzll(rest) + 2, 32u)"),
    # registers of the closed row N[u] (u itself included)
    ("twohop_open_rows", "feat",
     "        hll_add(r, (uint32_t)u);\n",
     "\n"),
    # the neighbours' registers are merged into the source's
    ("twohop_neighbours_not_merged", "feat",
     "for (int l = 0; l < 64; l++) r[l] = std::max(r[l], q[l]);",
     "(void)q;"),
    # linear counting for small estimates
    ("twohop_no_linear_counting", "feat",
     "if (e <= 160.0 && zeros > 0)",
     "if (e <= 0.0 && zeros > 0)"),
    # the register index is the hash's top 6 bits, not its lowest
    ("twohop_index_shares_rank_bits", "feat",
     "const uint32_t idx = (uint32_t)(h >> 58);",
     "const uint32_t idx = (uint32_t)((h << 6) >> 58);"),
]

MCE_PINS = ["test_bruteforce", "test_complete_graph", "test_complete_bipartite", "test_cycle",
            "test_moon_moser", "test_turan", "test_disjoint_union_and_isolated",
            "test_relabel_invariance_and_root_partition", "test_per_root_nodes_tiny",
            "test_per_root_nodes_tie_rule", "test_byte_model_hand_worked"]
FEAT_PINS = ["test_tiny_balls_near_exact", "test_star_and_complete", "test_bias_and_spread_on_skewed_graphs"]
GM_PINS = ["test_cliques_closed_form", "test_presets_vs_bruteforce",
           "test_random_patterns_vs_bruteforce", "test_kclique_independent_method",
           "test_per_root_counts_sum", "test_byte_model_hand_worked"]


def pin_cases(mod, names):
    """Every (label, function, kwargs) case of the named tests, parametrized
    ones expanded from their own parametrize marks."""
    for name in names:
        fn = getattr(mod, name)
        marks = [m for m in getattr(fn, "pytestmark", []) if m.name == "parametrize"]
        assert len(marks) <= 1, name
        if not marks:
            yield name, fn, {}
            continue
        keys = [k.strip() for k in marks[0].args[0].split(",")]
        for vals in marks[0].args[1]:
            vals = vals if isinstance(vals, tuple) else (vals,)
            yield f"{name}{vals}", fn, dict(zip(keys, vals))


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    with open(ORACLE_SRC) as f:
        src = f.read()
    cut = src.index(CHECKERS_MARK)
    out = tmp_path_factory.mktemp("oracle_mutants")
    procs = {}
    for name, _, old, new in MUTANTS:
        assert src.count(old) == 1, f"{name}: the original text must occur exactly once"
        assert src.index(old) < cut, f"{name}: mutants stay out of the independent checkers"
        cpp = out / f"{name}.cpp"
        cpp.write_text(src.replace(old, new))
        so = out / f"{name}.so"
        procs[name] = (so, subprocess.Popen(
            ["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-w", "-shared", "-o", str(so),
             str(cpp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    libs = {}
    for name, (so, p) in procs.items():
        log = p.communicate()[0].decode(errors="replace")
        assert p.returncode == 0, f"{name} failed to build:\n{log}"
        libs[name] = str(so)
    return libs


@pytest.mark.parametrize("name,suite", [(m[0], m[1]) for m in MUTANTS])
def test_pins_catch_planted_mistake(g2, orc, mutant_libs, monkeypatch, name, suite):
    monkeypatch.setattr(orc, "_lib", orc.bind(mutant_libs[name]))
    mod, names = {"mce": (TM, MCE_PINS), "gm": (TG, GM_PINS), "feat": (TF, FEAT_PINS)}[suite]
    caught = []
    for label, fn, kw in pin_cases(mod, names):
        try:
            fn(g2, orc, **kw)
        except AssertionError:
            caught.append(label)
            break
    assert caught, f"no pinning check fails with the planted mistake {name}"


def test_less_conditions_never_read():
    """The LESS branch of the condition check (gmatch.h:559-588) is not in
    MUTANTS because no plan the reference builds reaches it: over every preset
    and random fixture plan, LESS occurs only in the pre_cond rows of the root
    and of positions that are not shared, which the search never reads.  A mistake there cannot
    change a count, so no check could catch it."""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans.json")) as f:
        fx = json.load(f)
    plans = list(fx["presets"].values()) + [r["plan"] for r in fx["random"]]
    assert len(plans) == 72
    less_rows = 0
    for p in plans:
        for key in ("cond", "after_cond"):
            assert all(op != 0 for row in p[key] for op, _ in row), key
        for u, row in enumerate(p["pre_cond"]):
            if any(op == 0 for op, _ in row):
                less_rows += 1
                pos = p["rev"][u]
                assert pos == 0 or not p["shared"][pos]
    assert less_rows > 0


MLP_SRC = os.path.join(os.path.dirname(ORACLE_SRC), "mlp_ref.py")

# planted mistakes in the cost model's CPU reference (oracle/mlp_ref.py)
MLP_MUTANTS = [
    ("mlp_forward_drops_b2", "z2 = h1 @ r(W2) + b2", "z2 = h1 @ r(W2)"),
    ("mlp_backward_w2_not_transposed", "dh1 = (dh2 @ r(W2).T) * (z1 > 0)", "dh1 = (dh2 @ r(W2)) * (z1 > 0)"),
    ("mlp_backward_drops_layer2_relu_mask", "dh2 = np.outer(dy, r(W3)) * (z2 > 0)", "dh2 = np.outer(dy, r(W3))"),
    ("mlp_loss_gradient_not_averaged", "dy = 2 * e / n", "dy = 2 * e"),
    ("bf16_truncates", "u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16", "u = (u >> 16) << 16"),
    ("bf16_ties_away_from_zero", "u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16",
     "u = ((u + 0x8000) >> 16) << 16"),
]
MLP_PINS = ["test_fd_gradient", "test_bf16_rounding", "test_forward_and_gradient_match_torch",
            "test_bf16_rounding_matches_torch"]


@pytest.mark.parametrize("name,old,new", MLP_MUTANTS)
def test_mlp_pins_catch_planted_mistake(tmp_path, monkeypatch, name, old, new):
    with open(MLP_SRC) as f:
        src = f.read()
    assert src.count(old) == 1, f"{name}: the original text must occur exactly once"
    path = tmp_path / f"{name}.py"
    path.write_text(src.replace(old, new))
    spec = importlib.util.spec_from_file_location(f"mlp_ref_{name}", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    monkeypatch.setattr(TL, "mlp_ref", mod)
    caught = []
    for pin in MLP_PINS:
        try:
            getattr(TL, pin)()
        except AssertionError:
            caught.append(pin)
            break
    assert caught, f"no pinning check fails with the planted mistake {name}"
