"""The oracle's CSV input: this build's own format (the reference reads NDJSON
only), so these hand-written cases are what pins its specification: header
as a blank line, one object per record, null / number / string values,
quoting, and malformed records as the invalid line "{"."""
import pytest

import oracle

CASES = [
    # header only, with or without a newline
    (b"a,b", b"\n"),
    (b"a,b\n", b"\n"),
    # values: number, string, null (empty unquoted), trailing empty field
    (b"a,b,c\n1,x,\n", b'\n{"a":1,"b":"x","c":null}\n'),
    # JSON numbers stay numbers; other digit strings are strings
    (b"n\n-0.5e+3\n007\n1.\n-\n+1\n0\n", b'\n{"n":-0.5e+3}\n{"n":"007"}\n{"n":"1."}\n{"n":"-"}\n{"n":"+1"}\n{"n":0}\n'),
    # quoted fields are strings; "" is a quote; commas inside quotes
    (b'a,b\n"1","x,y"\n"say ""hi""",""\n', b'\n{"a":"1","b":"x,y"}\n{"a":"say \\"hi\\"","b":""}\n'),
    # CRLF, blank records stay blank lines
    (b"a\r\n1\r\n\r\n\n2\r\n", b'\n{"a":1}\n\n\n{"a":2}\n'),
    # escaping in strings: backslash, control bytes; UTF-8 kept
    (b'a\n"x\\y"\n"t\tb"\n\xc3\xa9\n', b'\n{"a":"x\\\\y"}\n{"a":"t\\u0009b"}\n{"a":"\xc3\xa9"}\n'),
    # malformed: wrong field count, quote in an unquoted field, text after a
    # closing quote, unterminated quote
    (b'a,b\n1\n1,2,3\nx"y,1\n"a"b,1\n"open,1\n1,2\n', b'\n{\n{\n{\n{\n{\n{"a":1,"b":2}\n'),
    # header names are JSON-escaped keys; duplicates kept (the last wins on lookup)
    (b'"k""q",k\\,k\n1,2,3\n', b'\n{"k\\"q":1,"k\\\\":2,"k":3}\n'),
    # true / false / null as text stay strings
    (b"v\ntrue\nnull\n", b'\n{"v":"true"}\n{"v":"null"}\n'),
    # a leading UTF-8 byte order mark is dropped; one elsewhere is data
    (b"\xef\xbb\xbfa,b\n1,\xef\xbb\xbf\n", b'\n{"a":1,"b":"\xef\xbb\xbf"}\n'),
]


@pytest.mark.parametrize("csv,want", CASES)
def test_csv_cases(csv, want):
    assert oracle.csv_to_ndjson(csv) == want


def test_csv_lines_keep_their_numbers():
    # one output line per input line, so strict-mode line numbers are the file's
    csv = b"a,b\n1,2\n\n3,4\nbad\n5,6"
    nd = oracle.csv_to_ndjson(csv)
    assert nd.count(b"\n") == csv.count(b"\n") + 1
    with pytest.raises(oracle.RunError) as e:
        oracle.run_csv("count", csv, on_invalid=oracle.ABORT)
    assert e.value.line == 5


def test_csv_queries_read_the_records():
    csv = b"host,status,bytes\na,200,10\nb,500,\nc,500,7\n"
    assert oracle.run_csv("select(.status == 500) | count", csv)[0] == b"2\n"
    assert oracle.run_csv("sum(.bytes)", csv)[0] == b"17\n"
    assert oracle.run_csv("select(.bytes == null) | .host", csv)[0] == b'"b"\n'


def test_csv_without_header_is_an_error():
    with pytest.raises(ValueError):
        oracle.csv_to_ndjson(b"")
    with pytest.raises(ValueError):
        oracle.csv_to_ndjson(b"\xef\xbb\xbf")
    with pytest.raises(ValueError):
        oracle.csv_to_ndjson(b'"open\n1\n')
