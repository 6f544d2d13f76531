"""Query front end (warpjq_amd/csrc/query.cpp) against the reference's
parser, lexer and lowering unit tests: crates/warpjq-core/src/query/mod.rs:
213-459, query/lex.rs:336-525 and gpu/lower.rs:189-265. Host code only."""
import ctypes

import pytest

import oracle
import warpjq_amd as w


def lowered(q, csv=False):
    c = w.Compiled(q, csv)
    v = c.view.contents
    blob = bytes(v.blob[i] for i in range(v.n_blob))
    steps = [(v.steps[i].kind, v.steps[i].index, blob[v.steps[i].key_off:v.steps[i].key_off + v.steps[i].key_len],
              v.steps[i].key_hash) for i in range(v.n_steps)]
    paths = [(v.paths[i].step_off, v.paths[i].step_count) for i in range(v.n_paths)]
    cmps = [dict(path=v.cmps[i].path, op=v.cmps[i].op, lit_kind=v.cmps[i].lit_kind,
                 lit=blob[v.cmps[i].lit_off:v.cmps[i].lit_off + v.cmps[i].lit_len], num=v.cmps[i].lit_num)
            for i in range(v.n_cmps)]
    rpn = [(v.cond_rpn[i].op, v.cond_rpn[i].arg) for i in range(v.n_cond)]
    prefixes = [blob[v.prefix_off[i]:v.prefix_off[i] + v.prefix_len[i]] for i in range(v.n_fields)]
    d = dict(steps=steps, paths=paths, cmps=cmps, rpn=rpn, blob=blob, n_needed=v.n_needed,
             needed=[v.needed_paths[i] for i in range(v.n_needed)],
             slot_of_path=[v.slot_of_path[i] for i in range(v.n_paths)], output_kind=v.output_kind,
             output_path=v.output_path, agg_kind=v.agg_kind, agg_path=v.agg_path, group_path=v.group_path,
             has_filter=v.has_filter, n_fields=v.n_fields, prefixes=prefixes,
             suffix=blob[v.suffix_off:v.suffix_off + v.suffix_len], csv_mode=v.csv_mode)
    d["_keep"] = c
    return d


def err(q):
    with pytest.raises(ValueError) as e:
        w.Compiled(q)
    return str(e.value)


CMP, TRUTHY, AND, OR, NOT = 0, 1, 2, 3, 4


def fnv1a(b):
    h = 0x811c9dc5
    for c in b:
        h = ((h ^ c) * 0x01000193) & 0xFFFFFFFF
    return h


def test_identity_and_paths():
    p = lowered(".")
    assert p["output_kind"] == 0 and p["has_filter"] == 0
    p = lowered(".a.b[2].c")
    assert [(s[0], s[1], s[2]) for s in p["steps"]] == [(0, 0, b"a"), (0, 0, b"b"), (1, 2, b""), (0, 0, b"c")]
    assert lowered('."content-type"')["steps"][0][2] == b"content-type"
    for k in (b"a", b"status", "é".encode()):
        assert fnv1a(k) == [s[3] for s in lowered('."%s"' % k.decode())["steps"]][0]


def test_select_count_and_conjunction():
    p = lowered("select(.status == 500) | count")
    assert p["has_filter"] == 1 and p["output_kind"] == 3 and p["agg_kind"] == 0 and p["agg_path"] == w.NONE
    assert p["cmps"][0]["num"] == 500.0 and p["rpn"] == [(CMP, 0)]
    assert [o for o, _ in lowered("select(.a == 1) | select(.b == 2)")["rpn"]] == [CMP, CMP, AND]
    # and binds tighter than or: a or (b and c)
    assert [o for o, _ in lowered("select(.a == 1 or .b == 2 and .c == 3)")["rpn"]] == [CMP, CMP, CMP, AND, OR]
    assert [o for o, _ in lowered("select(.ok | not)")["rpn"]] == [TRUTHY, NOT]


def test_projection_and_group_by():
    p = lowered("{time: .ts, msg}")
    assert p["output_kind"] == 2 and p["prefixes"] == [b'{"time":', b',"msg":'] and p["suffix"] == b"}"
    p = lowered("group_by(.host) | count")
    assert p["group_path"] != w.NONE and p["output_kind"] == 3


def test_interning_and_required_paths():
    assert len(lowered("select(.a == 1 or .a == 2) | {x: .a}")["paths"]) == 1
    assert lowered("select(.a == 1) | {x: .a, y: .b}")["n_needed"] == 2


def test_rejects_unsupported_by_name():
    for q in ["reduce", "map", "if", "def", "test", "length"]:
        m = err(q)
        assert "not in the v0.1 subset" in m or "roadmap" in m, m
    for q, want in [("reduce .[] as $x (0; .+$x)", "`reduce` is not"), ("def f: .; f", "user-defined functions"),
                    ("if .a then 1 else 2 end", "`if/then/else` is not"), ("foreach .[] as $x (0; .+1)", "`foreach` is not")]:
        assert want in err(q)
    for q in ["select(.a == 1) | map(.b)", '.a | test("x")', "select(.a == 1) | keys"]:
        m = err(q)
        assert "not in the v0.1 subset" in m or "roadmap" in m


def test_pipes_in_strings_and_parens():
    w.Compiled('select(.msg == "a|b")')
    w.Compiled("select(.a == 1 or (.b == 2))")
    w.Compiled("select(.a | not)")


def test_nesting_limit():
    for depth in (100, 1000, 10000, 100000):
        m = err("select(" + "(" * depth)
        assert "nest more than" in m or "expected" in m
    err("select(" + "(" * 500 + ".a == 1" + ")" * 500 + ")")
    for depth in (1, 8, 32, 64):
        w.Compiled("select(" + "(" * depth + ".a == 1" + ")" * depth + ")")


def test_hints_and_rejections():
    assert "did you mean `==`" in err("select(.a = 1)")
    assert "did you mean `.status`?" in err("select(status == 500)")
    err("group_by(.host)")
    err("group_by(.host) | {a: .b}")
    assert "field-to-field" in err("select(.a == .b)")
    err("count | count")
    err("{a: .b} | {c: .d}")


def test_rpn_depth_and_blob():
    p = lowered("select((.a == 1 and .b == 2) or .c == 3)")
    assert len(p["cmps"]) == 3 and [o for o, _ in p["rpn"]] == [CMP, CMP, AND, CMP, OR]
    c = lowered('select(.method == "POST")')["cmps"][0]
    assert c["lit_kind"] == 4 and c["lit"] == b"POST"


def test_lexer_numbers_and_strings():
    # lex.rs:371-437 through the literal each produces
    for text, num in [("42", 42.0), ("-42", -42.0), ("0", 0.0), ("4.5", 4.5), ("1e3", 1000.0), ("1E3", 1000.0),
                      ("1e+3", 1000.0), ("1e-3", 0.001), ("-2.5E+10", -2.5e10), ("1.", 1.0),
                      ("99999999999999999999", 1e20)]:
        assert lowered(f"select(.a == {text})")["cmps"][0]["num"] == num
    for m, q in [("expected a number", "select(.a == -)"), ("no digits in its exponent", "select(.a == 1e)"),
                 ("no digits in its exponent", "select(.a == 1e+)"), ("no digits in its exponent", "select(.a == 1E-)")]:
        assert m in err(q)
    for src, lit in [(r'"plain"', b"plain"), (r'"a\"b"', b'a"b'), (r'"a\\b"', b"a\\b"), (r'"a\/b"', b"a/b"),
                     (r'"\b\f\n\r\t"', b"\x08\x0c\n\r\t"), (r'"A"', b"A"), (r'"é"', "é".encode()),
                     (r'"日"', "日".encode()), (r'"😀"', "😀".encode()), ('""', b""),
                     ('"日本"', "日本".encode()), (r'"\ud83d"', "�".encode()),
                     (r'"\ud83dx"', "�x".encode()), (r'"\ud83dA"', "�A".encode())]:
        assert lowered(f"select(.a == {src})")["cmps"][0]["lit"] == lit, src
    assert "unterminated" in err('select(.a == "unterminated)')
    assert "unknown escape" in err(r'select(.a == "bad \q escape")')
    assert "`\\u`" in err(r'select(.a == "\u00")')
    assert "`\\u`" in err(r'select(.a == "\uZZZZ")')
    assert "unexpected character `@`" in err("@")
    assert "`$`" in err("$x")
    assert "expected `!=`" in err("!")
    w.Compiled(".a # trailing comment")
    w.Compiled("# leading\n.a")


def test_lowering_golden():
    # gpu/lower.rs:189-265
    p = lowered("select(.a == 1) | {x: .b}")
    assert p["n_needed"] == 2 and all(s != w.NONE for s in p["slot_of_path"])
    p = lowered("{a: .x, b: .y}")
    assert p["prefixes"] == [b'{"a":', b',"b":'] and p["suffix"] == b"}"
    assert lowered(r'{"a\"b": .x}')["prefixes"][0] == b'{"a\\"b":'
    p = lowered("{a: .x, b: .y}", csv=True)
    assert p["prefixes"] == [b"", b","] and p["suffix"] == b"" and p["csv_mode"] == 1
    p = lowered("count")
    assert (p["agg_kind"], p["agg_path"], p["group_path"]) == (0, w.NONE, w.NONE)
    p = lowered("group_by(.h) | sum(.b)")
    assert p["agg_kind"] == 1 and p["agg_path"] != w.NONE and p["group_path"] != w.NONE
    assert lowered(".")["has_filter"] == 0


def test_csv_headers_through_oracle():
    # ir.rs:237-251 header rules, rendered by the writer
    data = b'{"a":{"b":[1]},"h":"x"}\n'
    assert oracle.run(".a.b[0]", data, fmt=oracle.FMT_CSV)[0] == b".a.b[0]\n1\n"
    assert oracle.run('."odd key"', b'{"odd key":1}\n', fmt=oracle.FMT_CSV)[0] == b'".""odd key"""\n1\n'
    assert oracle.run("group_by(.h) | count", data, fmt=oracle.FMT_CSV)[0] == b".h,count\nx,1\n"
