"""Parser for the subset of HOCON used by the reference's configs
(config/diligent.conf, config/own_objects.conf): nested ``name { ... }``
blocks, ``key = value`` / ``key: value``, numbers, booleans, bare or quoted
strings, ``[a, b]`` lists, ``#`` and ``//`` comments.  Returns a ``Conf``
(dict) that also answers dotted lookups like ``conf['model.sdf_network']``
and ``get_int/get_float/get_bool`` with defaults, as pyhocon's ConfigTree does
for exp_runner.py:45-83.
"""
from __future__ import annotations

import re


class Conf(dict):
    def _lookup(self, key):
        node = self
        for part in key.split("."):
            if not isinstance(node, dict) or part not in node:
                raise KeyError(key)
            node = node[part]
        return node

    def __getitem__(self, key):
        if isinstance(key, str) and "." in key and not dict.__contains__(self, key):
            return self._lookup(key)
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def get_int(self, key, default=None):
        v = self.get(key, default)
        return None if v is None else int(v)

    def get_float(self, key, default=None):
        v = self.get(key, default)
        return None if v is None else float(v)

    def get_bool(self, key, default=None):
        v = self.get(key, default)
        if isinstance(v, str):
            return v.lower() == "true"
        return None if v is None else bool(v)


_TOKEN = re.compile(r'[ \t\r]*(?:(#|//)[^\n]*|("(?:[^"\\]|\\.)*")|([{}\[\],=:\n])|([^\s{}\[\],=:#"]+))')


def _tokens(text):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            if text[pos].isspace():
                pos += 1
                continue
            raise ValueError(f"HOCON parse error at {text[pos:pos + 30]!r}")
        pos = m.end()
        if m.group(1):
            continue
        if m.group(2):
            out.append(("str", m.group(2)[1:-1]))
        elif m.group(3):
            if m.group(3) != "\n":
                out.append(("sym", m.group(3)))
            else:
                out.append(("nl", "\n"))
        elif m.group(4):
            out.append(("word", m.group(4)))
    return out


def _scalar(word):
    low = word.lower()
    if low == "true":
        return True
    if low == "false":
        return False
    if low in ("null", "none"):
        return None
    try:
        if re.fullmatch(r"[+-]?\d+", word):
            return int(word)
        return float(word)
    except ValueError:
        return word


def parse(text: str) -> Conf:
    toks = _tokens(text)
    i = 0

    def skip_nl():
        nonlocal i
        while i < len(toks) and toks[i][0] == "nl":
            i += 1

    def value():
        nonlocal i
        skip_nl()
        kind, v = toks[i]
        if kind == "sym" and v == "{":
            i += 1
            return obj(end="}")
        if kind == "sym" and v == "[":
            i += 1
            items = []
            while True:
                skip_nl()
                if toks[i] == ("sym", "]"):
                    i += 1
                    return items
                items.append(value())
                skip_nl()
                if toks[i] == ("sym", ","):
                    i += 1
        i += 1
        if kind == "str":
            return v
        # bare words up to the end of the line form one value (e.g. paths)
        words = [v]
        while i < len(toks) and toks[i][0] == "word":
            words.append(toks[i][1])
            i += 1
        return _scalar(" ".join(words)) if len(words) > 1 else _scalar(v)

    def obj(end=None):
        nonlocal i
        out = Conf()
        while True:
            skip_nl()
            if i >= len(toks):
                if end:
                    raise ValueError("HOCON: unterminated block")
                return out
            if end and toks[i] == ("sym", end):
                i += 1
                return out
            if toks[i] == ("sym", ","):
                i += 1
                continue
            kind, key = toks[i]
            i += 1
            skip_nl()
            if toks[i] in (("sym", "="), ("sym", ":")):
                i += 1
            v = value()
            node = out
            parts = key.split(".")
            for part in parts[:-1]:
                node = node.setdefault(part, Conf())
            if isinstance(v, dict) and isinstance(node.get(parts[-1]), dict):
                node[parts[-1]].update(v)
            else:
                node[parts[-1]] = v

    return obj()
