"""Minimal TOML-subset reader for the test harness.

Supports what pic-nix ``config.toml`` files use: ``[table]``, ``[a.b]``,
``[[array.of.tables]]``, ``key = value`` with integers, floats, booleans,
single/double quoted strings, flat inline arrays, and ``#`` comments.
"""


def _value(tok):
    tok = tok.strip()
    if tok.startswith("[") and tok.endswith("]"):
        inner = tok[1:-1].strip()
        return [] if not inner else [_value(t) for t in inner.split(",") if t.strip()]
    if tok[:1] in ("'", '"'):
        return tok[1:-1]
    if tok in ("true", "false"):
        return tok == "true"
    try:
        return int(tok)
    except ValueError:
        return float(tok)


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            if ch == q:
                q = None
        elif ch in ("'", '"'):
            q = ch
        elif ch == "#":
            break
        out.append(ch)
    return "".join(out).strip()


def loads(text):
    root = {}
    cur = root
    for raw in text.splitlines():
        line = _strip_comment(raw)
        if not line:
            continue
        if line.startswith("[["):
            path = line[2:-2].strip().split(".")
            node = root
            for k in path[:-1]:
                node = node.setdefault(k, {})
                if isinstance(node, list):
                    node = node[-1]
            lst = node.setdefault(path[-1], [])
            lst.append({})
            cur = lst[-1]
        elif line.startswith("["):
            path = line[1:-1].strip().split(".")
            node = root
            for k in path:
                node = node.setdefault(k, {})
                if isinstance(node, list):
                    node = node[-1]
            cur = node
        else:
            key, val = line.split("=", 1)
            cur[key.strip()] = _value(val)
    return root


def load(path):
    with open(path) as f:
        return loads(f.read())
