"""The process boundary on the CPU: the case-document parser and its
STRUT_JSON_CACHE_V1 file against the reference's own test vectors
(tests/mojo/test_json_native.mojo, tests/mojo/test_json_cache.mojo) and a
cache text written out by hand from the format (strut_io.mojo:129-179),
the command line (strut.mojo:13-82), the time-series file scanner
(input_types.mojo:1603-1676) and the recorder files
(transient_nonlinear.mojo:1553-1695, 3228-3432) from a synthetic history.
Nothing here touches the GPU."""

import json

import numpy as np
import pytest

from strut_amd import cli
from strut_amd.json_native import (ARRAY, BOOL, NULL, NUMBER, OBJECT, STRING, CacheError, CaseSourceInfo,
                                   JsonParseError, load_case_document, load_json_cache, parse_json_native,
                                   serialize_json_cache, write_json_cache)


def test_parse_basic_object_and_array():
    doc = parse_json_native('{"flag":true,"count":12,"name":"beam","items":[null,false,1.25e3]}')
    root = doc.root_index
    assert doc.nodes[root].tag == OBJECT
    flag, count, name, items = (doc.object_find(root, k) for k in ("flag", "count", "name", "items"))
    assert doc.nodes[flag].bool_value is True
    assert doc.nodes[count].number_value == 12.0 and doc.nodes[count].number_is_integer
    assert doc.nodes[name].text_value == "beam"
    assert doc.nodes[items].tag == ARRAY and doc.nodes[items].child_count == 3
    assert doc.nodes[doc.array_item(items, 0)].tag == NULL
    assert doc.nodes[doc.array_item(items, 1)].bool_value is False
    third = doc.nodes[doc.array_item(items, 2)]
    assert third.number_value == 1250.0 and not third.number_is_integer and third.text_value == "1.25e3"
    # post-order: children first, root last; byte spans of the source text
    assert [n.tag for n in doc.nodes] == [BOOL, NUMBER, STRING, NULL, BOOL, NUMBER, ARRAY, OBJECT]
    assert root == 7 and (doc.nodes[root].start, doc.nodes[root].end) == (0, 66)
    assert (doc.nodes[name].start, doc.nodes[name].end) == (31, 37)
    assert doc.object_entries == [("flag", 0, 1), ("count", 1, 13), ("name", 2, 24), ("items", 6, 38)]


def test_parse_string_escapes_and_unicode():
    doc = parse_json_native('{"text":"line\\nquote: \\"x\\" slash: \\\\ greek: \\u03B1 fire: \\uD83D\\uDD25"}')
    assert doc.to_python()["text"] == "line\nquote: \"x\" slash: \\ greek: α fire: \U0001F525"


@pytest.mark.parametrize("text,message", [
    ("{\n  \"items\": [1, 2,],\n  \"ok\": true\n}", "JSON parse error at line 2, column 18"),
    ('{"items":[1,2,]}', "trailing comma in array"),
    ('{"id":01}', "numbers cannot have leading zeroes"),
    ('{"bad":"\\q"}', "invalid escape sequence"),
    ('{"ok":true} []', "unexpected trailing content"),
    ('{"a":1,}', "trailing comma in object"),
    ('[1.]', "fractional number is missing digits"),
    ('[1e+]', "exponent is missing digits"),
    ('["\\ud800x"]', "high surrogate must be followed by low surrogate"),
    ('["\\udc00"]', "low surrogate without preceding high surrogate"),
])
def test_parse_errors_report_the_reference_messages(text, message):
    with pytest.raises(JsonParseError, match=message.replace("[", r"\[")):
        parse_json_native(text)


def test_parse_error_rendering_has_context_and_caret():
    with pytest.raises(JsonParseError) as e:
        parse_json_native('{\n  "a": [1, 2,],\n}')
    assert str(e.value) == 'JSON parse error at line 2, column 14: trailing comma in array\n  "a": [1, 2,],\n' + \
        " " * 13 + "^"


def test_first_duplicate_key_wins_as_object_find_does():
    doc = parse_json_native('{"a":1,"a":2}')
    assert doc.object_find(doc.root_index, "a") == 0 and doc.to_python() == {"a": 1}


CACHE_TEXT = ("STRUT_JSON_CACHE_V1\n11\n/tmp/c.json\n4\n/tmp\n3\n4\n"
              "2\t6\t7\t0\t1\t0\t0\t1\n1\n"
              "3\t8\t11\t0\t0\t0\t0\t1\nx\n"
              "4\t5\t12\t0\t0\t0\t2\t0\n\n"
              "5\t0\t13\t0\t0\t0\t1\t0\n\n"
              "1\n2\t1\t1\na\n"
              "2\n0\n1\n")


def test_cache_bytes_follow_the_format():
    doc = parse_json_native('{"a":[1,"x"]}')
    assert serialize_json_cache(doc, CaseSourceInfo("/tmp/c.json", "/tmp")).decode() == CACHE_TEXT


def test_cache_round_trips_document_and_source_info(tmp_path):
    doc = parse_json_native('{"model":{"ndm":2,"ndf":3},"items":[true,1.25,"beam"]}')
    path = tmp_path / "case.bin"
    info = CaseSourceInfo("/tmp/example_case.json", "/tmp")
    write_json_cache(str(path), doc, info)
    loaded, linfo = load_json_cache(str(path))
    assert (linfo.json_path, linfo.case_dir) == (info.json_path, info.case_dir)
    assert loaded.root_index == doc.root_index
    assert loaded.nodes == doc.nodes and loaded.object_entries == doc.object_entries
    assert loaded.array_items == doc.array_items
    items = loaded.object_find(loaded.root_index, "items")
    assert loaded.nodes[items].tag == ARRAY
    assert loaded.nodes[loaded.array_item(items, 1)].number_value == 1.25
    assert loaded.to_python() == {"model": {"ndm": 2, "ndf": 3}, "items": [True, 1.25, "beam"]}


def test_cache_with_multiline_and_unicode_strings(tmp_path):
    doc = parse_json_native('{"k\\né":"two\\nlines α","n":-0.0}')
    path = tmp_path / "c.bin"
    write_json_cache(str(path), doc, CaseSourceInfo("", ""))
    loaded, info = load_json_cache(str(path))
    assert loaded.to_python() == doc.to_python()
    assert info.json_path == str(path.resolve())  # an empty recorded path falls back to the cache's own


@pytest.mark.parametrize("text,message", [
    ("STRUT_JSON_CACHE_V2\n", "magic"),
    ("STRUT_JSON_CACHE_V1\n5\nab", "truncated block"),
    ("STRUT_JSON_CACHE_V1\n2\nabc\n", "missing block terminator"),
    ("STRUT_JSON_CACHE_V1\n0\n\n0\n\n3", "truncated line"),
])
def test_cache_errors(tmp_path, text, message):
    p = tmp_path / "bad.bin"
    p.write_text(text)
    with pytest.raises(CacheError, match=message):
        load_json_cache(str(p))


def test_load_case_document_sources(tmp_path):
    case = tmp_path / "case.json"
    case.write_text('{"x": [1, 2.5]}')
    with pytest.raises(ValueError, match="use only one"):
        load_case_document(str(case), str(case), "")
    with pytest.raises(ValueError, match="missing --input"):
        load_case_document("", "", "")
    doc, info = load_case_document(str(case), "", str(tmp_path / "c.bin"))
    assert info.case_dir == str(tmp_path.resolve())
    doc2, info2 = load_case_document("", str(tmp_path / "c.bin"), "")
    assert doc2.to_python() == doc.to_python() == {"x": [1, 2.5]} and info2.json_path == info.json_path


def test_command_line():
    a = cli.parse_args(["strut", "--input", "c.json", "--output", "out", "--compute-only", "--profile", "p.json"])
    assert (a.input, a.output, a.profile, a.compute_only, a.batch, a.input_bin) == \
        ("c.json", "out", "p.json", True, "", "")
    assert cli.parse_args(["strut", "--output"]).output == ""  # a flag with no word after it is absent
    assert cli.parse_args(["strut", "--input", "a", "--input", "b"]).input == "a"


def test_write_input_bin_without_output_writes_the_cache_only(tmp_path, capsys):
    case = tmp_path / "case.json"
    case.write_text(json.dumps({"analysis": {"type": "static_linear"}}))
    assert cli.main(["strut", "--input", str(case), "--write-input-bin", str(tmp_path / "c.bin")]) == 0
    assert (tmp_path / "c.bin").read_bytes().startswith(b"STRUT_JSON_CACHE_V1\n")
    assert cli.main(["strut", "--input", str(case)]) == 2
    assert "usage:" in capsys.readouterr().err


def test_cases_outside_the_device_path_are_refused(tmp_path, capsys):
    case = tmp_path / "case.json"
    case.write_text(json.dumps({"model": {"ndm": 2, "ndf": 3}, "nodes": [], "analysis": {"type": "static_linear"}}))
    assert cli.main(["strut", "--input", str(case), "--output", str(tmp_path / "o")]) == 2
    assert "outside the device path" in capsys.readouterr().err
    assert not (tmp_path / "o" / "analysis_time_us.txt").exists()


def test_numeric_values_text():
    text = "1.0 2.5e-3\n-4D2, +.5 x7 1e 5. -.e .25E+1 3d-1\t+-8"
    assert cli.parse_numeric_values_text(text, "f") == [1.0, 2.5e-3, -400.0, 0.5, 7.0, 1.0, 5.0, 2.5, 0.3, -8.0]
    with pytest.raises(ValueError, match="no numeric values"):
        cli.parse_numeric_values_text("abc -. e", "f")


def test_time_series_files_are_resolved_against_the_case_directory(tmp_path):
    (tmp_path / "gm.txt").write_text("0.0\n0.1\n-0.2\n")
    case = {"time_series": [{"type": "PathFile", "tag": 1, "dt": 0.01, "values_path": "gm.txt"}]}
    cli.resolve_time_series(case, CaseSourceInfo(str(tmp_path / "c.json"), str(tmp_path)))
    assert case["time_series"][0]["values"] == [0.0, 0.1, -0.2] and case["time_series"][0]["type"] == "Path"
    with pytest.raises(ValueError, match="cannot specify both"):
        cli.resolve_time_series({"time_series": [{"type": "Path", "values": [1], "path": "x"}]}, CaseSourceInfo())
    with pytest.raises(ValueError, match="not found"):
        cli.resolve_time_series({"time_series": [{"type": "Path", "path": "missing.txt"}]}, CaseSourceInfo())


def _recorder_case():
    nodes = [{"id": 10, "x": 0.0, "y": 0.0}, {"id": 20, "x": 0.0, "y": 3.0}, {"id": 30, "x": 4.0, "y": 3.0}]
    recs = [{"type": "node_displacement", "nodes": [20], "dofs": [1, 3], "output": "roof"},
            {"type": "drift", "i_node": 10, "j_node": 20, "dof": 1, "perp_dirn": 2, "output": "d1"},
            {"type": "envelope_node_displacement", "nodes": [20, 30], "dofs": [1]},
            {"type": "node_displacement", "nodes": [20], "dofs": [2], "output": "roof"},
            {"type": "node_displacement", "nodes": [30], "dofs": [1]}]
    return {"nodes": nodes, "recorders": recs}


def test_recorder_files_from_a_history():
    plan = cli.recorder_plan(_recorder_case(), ndf=3, has_z=False)
    # DOF k of node index i is column 3 i + k - 1
    assert plan.columns == [3, 5, 0, 6, 4]
    assert plan.files == ["roof_node20.out", "d1_i10_j20.out", "envelope_node_displacement_node20.out",
                          "envelope_node_displacement_node30.out", "node_disp_node30.out"]
    hist = np.array([[0.1, -2.5, 0.0, 1e-05, 7.0],
                     [-0.3, 1.0, 0.03, -0.0, 2.0],
                     [0.2, 0.5, -0.03, 0.25, -9.0]])
    out = cli.recorder_outputs(plan, hist)
    # two recorders on one file: their lines interleave per step in recorder order
    assert out["roof_node20.out"] == "0.1 -2.5\n7.0\n-0.3 1.0\n2.0\n0.2 0.5\n-9.0\n"
    assert out["d1_i10_j20.out"] == "".join(repr(v) + "\n" for v in [(0.1 - 0.0) / 3.0, (-0.3 - 0.03) / 3.0,
                                                                     (0.2 + 0.03) / 3.0])
    assert out["envelope_node_displacement_node20.out"] == "-0.3\n0.2\n0.3\n"
    # -0.0 after 1e-05: the first value is kept on ties, as the strict comparisons do
    assert out["envelope_node_displacement_node30.out"] == "-0.0\n0.25\n0.25\n"
    assert out["node_disp_node30.out"] == "1e-05\n-0.0\n0.25\n"


def test_recorder_plan_refusals():
    case = _recorder_case()
    case["recorders"] = [{"type": "drift", "i_node": 10, "j_node": 20, "dof": 1, "perp_dirn": 3}]
    with pytest.raises(ValueError, match="requires z"):
        cli.recorder_plan(case, ndf=3, has_z=False)
    case["recorders"] = [{"type": "drift", "i_node": 20, "j_node": 30, "dof": 1, "perp_dirn": 2}]
    with pytest.raises(ValueError, match="denominator is zero"):
        cli.recorder_plan(case, ndf=3, has_z=False)
    case["recorders"] = [{"type": "element_force", "elements": [1]}]
    with pytest.raises(cli.CaseOutsidePath):
        cli.recorder_plan(case, ndf=3, has_z=False)
    case["recorders"] = [{"type": "node_displacement", "nodes": [99], "dofs": [1]}]
    with pytest.raises(ValueError, match="recorder node not found"):
        cli.recorder_plan(case, ndf=3, has_z=False)


def test_phase_times_layout():
    text = cli.phase_times_json({"analysis_us": 12, "global_nonlinear_iterations": 7}, {},
                                {"forceBeamColumn2d": 21, "dispBeamColumn2d": 0}, [3, 4])
    d = json.loads(text)
    assert list(d)[:len(cli.PHASE_KEYS)] == list(cli.PHASE_KEYS)
    assert d["analysis_us"] == 12 and d["global_nonlinear_iterations"] == 7
    assert d["element_type_timing_us"] == {} and d["element_type_call_counts"] == {"forceBeamColumn2d": 21}
    assert d["step_newton_iterations"] == [3, 4]
