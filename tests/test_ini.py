"""Settings.ini semantics of the reference's vendored inih + INIReader
(F/INIlib/ini.c, F/INIlib/INIReader.cpp), through fdtd2d_ini_lookup."""

import pytest

from fdtd2d_amd._lib import ini_lookup

INT, REAL, BOOL = 0, 1, 2


def look(text, sec, key, kind=REAL, default=-7.0):
    return ini_lookup(text, sec, key, kind, default)


def test_basic_keys_and_inline_comments():
    t = "[grid]\nImax = 250\t\t; grid size (mm)\nds = 0.5 ; spacing\n"
    assert look(t, "grid", "Imax", INT) == (250, 0)
    assert look(t, "grid", "ds") == (0.5, 0)


def test_semicolon_without_preceding_space_is_not_a_comment():
    t = "[a]\nx=12;5\n"
    # value is "12;5": strtol reads 12
    assert look(t, "a", "x", INT)[0] == 12


def test_case_insensitive_sections_and_keys():
    t = "[PulseParam]\nFreqA=38\n"
    assert look(t, "pulseparam", "FREQA")[0] == 38.0


def test_strtol_base_detection():
    t = "[g]\nh=0x20\no=010\nf=0.5\nneg=-3\n"
    assert look(t, "g", "h", INT)[0] == 32
    assert look(t, "g", "o", INT)[0] == 8
    assert look(t, "g", "f", INT)[0] == 0  # "0.5" -> 0 as an integer
    assert look(t, "g", "neg", INT)[0] == -3


def test_missing_and_unparsable_fall_back_to_default():
    t = "[g]\nx=abc\n"
    assert look(t, "g", "x", INT, 41)[0] == 41
    assert look(t, "g", "nope", REAL, 2.5)[0] == 2.5
    assert look(t, "other", "x", REAL, 1.5)[0] == 1.5


def test_booleans():
    t = "[p]\na=yes\nb=Off\nc=1\nd=maybe\n"
    assert look(t, "p", "a", BOOL, 0)[0] == 1
    assert look(t, "p", "b", BOOL, 1)[0] == 0
    assert look(t, "p", "c", BOOL, 0)[0] == 1
    assert look(t, "p", "d", BOOL, 1)[0] == 1  # default


def test_duplicate_keys_join_and_integer_reads_the_first():
    t = "[g]\nx=5\nx=9\n"
    assert look(t, "g", "x", INT)[0] == 5


def test_continuation_line_joins_previous_value():
    t = "[g]\nx=7\n  8\ny=1\n"
    assert look(t, "g", "x", INT)[0] == 7
    assert look(t, "g", "y", INT)[0] == 1


def test_comment_lines_and_bom_and_crlf():
    t = "﻿; header\r\n# another\r\n[g]\r\nx = 3\r\n"
    assert look(t.encode("utf-8").decode("latin-1"), "g", "x", INT) == (3, 0)


def test_parse_error_records_first_bad_line_and_continues():
    t = "[g]\nx=1\nthis line has no equals\n[h\ny=2\n"
    v, st = look(t, "g", "y", INT)
    assert st == 3
    assert look(t, "g", "x", INT)[0] == 1


def test_long_lines_split_at_199_characters():
    long_value = "1" * 250
    t = "[g]\nx=" + long_value + "\n"
    # the first 199 characters of the line are one line: "x=" + 197 ones
    v, st = look(t, "g", "x", REAL)
    assert v == pytest.approx(float("1" * 197), rel=1e-12)
    assert st == 3  # the remainder has no '=' -> parse error on line 3


def test_colon_separator():
    t = "[g]\nx: 4\n"
    assert look(t, "g", "x", INT)[0] == 4
