"""Declarative `.case` files: a schema of typed keys and the child-process runner.

A case file is INI text.  A :class:`Schema` lists every accepted
``section.key`` once, as a :class:`Key` row: its default, the native flag it
becomes (if any) and a value type that checks and normalises the text.
Cross-key constraints are plain functions over the resolved values.  The H8
launcher (``nodals_amd.case``) and the k-omega SST launcher
(``nodals_amd.sst_case``) are two schemas over this one engine.

:func:`run_logged` starts the native solver as a child process in its own
process group, copies its output line by line to the terminal (optionally)
and to a log file from a reader thread, and enforces the time limit while
the child is still printing (exit status 124 on expiry).
"""
from __future__ import annotations

import configparser
import os
import signal
import subprocess
import sys
import threading
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Dict, Iterable, List, Mapping, Optional, Sequence, Tuple


class CaseError(ValueError):
    """A case file that cannot be translated (the launchers exit with 4)."""


# ---------------------------------------------------------------- value types
# A value type maps (raw text, "[section] key") to the canonical text, or
# raises CaseError.


def text(raw: str, where: str) -> str:
    return raw.strip()


def path_text(raw: str, where: str) -> str:
    return os.path.expandvars(os.path.expanduser(raw.strip()))


def integer(raw: str, where: str) -> str:
    try:
        return str(int(raw.strip()))
    except ValueError:
        raise CaseError(f"{where}: expected an integer, got {raw!r}") from None


def number(raw: str, where: str) -> str:
    s = raw.strip()
    try:
        float(s)
    except ValueError:
        raise CaseError(f"{where}: expected a number, got {raw!r}") from None
    return s


_TRUTH = {"1": True, "true": True, "yes": True, "on": True, "0": False, "false": False, "no": False, "off": False}


def boolean(raw: str, where: str) -> str:
    v = _TRUTH.get(raw.strip().lower())
    if v is None:
        raise CaseError(f"{where}: expected boolean, got {raw!r}")
    return "1" if v else "0"


def one_of(*allowed: str, message: Optional[str] = None):
    """Lower-cased choice among `allowed`; `message` replaces the default
    complaint (launchers keep the reference's wording)."""
    allowed_set = frozenset(allowed)

    def check(raw: str, where: str) -> str:
        v = raw.strip().lower()
        if v not in allowed_set:
            raise CaseError(message or f"{where} must be one of {', '.join(allowed)}")
        return v

    return check


def only(value: str, message: str):
    """A key this build accepts with a single value."""

    def check(raw: str, where: str) -> str:
        v = raw.strip().lower()
        if v != value:
            raise CaseError(message)
        return v

    return check


# ---------------------------------------------------------------- schema rows


@dataclass(frozen=True)
class Key:
    """One accepted case key.

    name    "section.key"
    default canonical text used when the case omits the key (None: required
            if `flag` is set, else simply absent)
    flag    native command-line flag it is emitted as (None: consumed by
            rules only)
    kind    value type (see above)
    when    emit the flag only if this predicate of the resolved values holds
    """

    name: str
    default: Optional[str] = None
    flag: Optional[str] = None
    kind: Callable[[str, str], str] = text
    when: Optional[Callable[[Mapping[str, str]], bool]] = None

    @property
    def section(self) -> str:
        return self.name.split(".", 1)[0]

    @property
    def key(self) -> str:
        return self.name.split(".", 1)[1]

    @property
    def where(self) -> str:
        return f"[{self.section}] {self.key}"


Rule = Callable[[Mapping[str, str]], None]


@dataclass
class Schema:
    """An ordered key table plus cross-key rules.

    `label` appears in the unknown-section/key messages ("GPU case" for H8).
    `required` names keys that must be present and non-empty.
    """

    label: str
    keys: Sequence[Key]
    rules: Sequence[Rule] = ()
    required: Sequence[str] = ()
    _index: Dict[str, Key] = field(init=False, repr=False)

    def __post_init__(self) -> None:
        self._index = {k.name: k for k in self.keys}
        if len(self._index) != len(self.keys):
            raise AssertionError("duplicate schema key")

    @property
    def sections(self) -> frozenset:
        return frozenset(k.section for k in self.keys)

    def key(self, section: str, key: str) -> Key:
        if section not in self.sections:
            raise CaseError(f"unknown {self.label} section [{section}]")
        k = self._index.get(f"{section}.{key}")
        if k is None:
            raise CaseError(f"unknown {self.label} key [{section}] {key}")
        return k

    def check_names(self, doc: "CaseDoc") -> None:
        for section, key in doc.names():
            self.key(section, key)

    def resolve(self, doc: "CaseDoc") -> Dict[str, str]:
        """Canonical value of every key (defaults filled in), after checking
        names, required keys, value types and rules."""
        self.check_names(doc)
        vals: Dict[str, str] = {}
        for k in self.keys:
            raw = doc.raw(k.section, k.key)
            if raw is None or (raw.strip() == "" and k.name in self.required):
                if k.name in self.required:
                    raise CaseError(f"missing required {k.where}")
                if k.default is None:
                    continue
                raw = k.default
            vals[k.name] = k.kind(raw, k.where)
        for rule in self.rules:
            rule(vals)
        return vals

    def argv(self, doc: "CaseDoc") -> List[str]:
        vals = self.resolve(doc)
        out: List[str] = []
        for k in self.keys:
            if k.flag is None or k.name not in vals:
                continue
            if k.when is not None and not k.when(vals):
                continue
            out += [k.flag, vals[k.name]]
        return out


# ---------------------------------------------------------------- documents


class CaseDoc:
    """The (section, key) -> raw text content of one case file, with
    ``--set section.key=value`` overrides applied on top."""

    def __init__(self, data: Optional[Dict[Tuple[str, str], str]] = None, source: Optional[Path] = None):
        self._data: Dict[Tuple[str, str], str] = dict(data or {})
        self.source = source

    @classmethod
    def read(cls, path: Path) -> "CaseDoc":
        p = Path(path)
        cp = configparser.ConfigParser(interpolation=None, inline_comment_prefixes=("#", ";"), strict=True)
        cp.optionxform = str.lower  # type: ignore[assignment]
        try:
            with p.open("r", encoding="utf-8") as f:
                cp.read_file(f)
        except configparser.Error as e:
            raise CaseError(f"{p}: {e}") from None
        data = {(s.lower(), k): v for s in cp.sections() for k, v in cp.items(s)}
        return cls(data, source=p)

    def names(self) -> List[Tuple[str, str]]:
        return list(self._data)

    def raw(self, section: str, key: str) -> Optional[str]:
        return self._data.get((section, key))

    def get(self, section: str, key: str, default: Optional[str] = None) -> Optional[str]:
        v = self._data.get((section, key))
        return default if v is None else v.strip()

    def set(self, section: str, key: str, value: str) -> None:
        self._data[(section.strip().lower(), key.strip().lower())] = str(value)

    def apply(self, schema: Schema, overrides: Iterable[str]) -> None:
        """Apply ``section.key=value`` items; names are checked against the
        schema before anything is changed."""
        parsed = []
        for item in overrides:
            name, sep, value = item.partition("=")
            section, dot, key = name.partition(".")
            if not sep or not dot or not section.strip() or not key.strip():
                raise CaseError(f"--set expects section.key=value, got {item!r}")
            k = schema.key(section.strip().lower(), key.strip().lower())
            parsed.append((k.section, k.key, value))
        for section, key, value in parsed:
            self.set(section, key, value)


# ---------------------------------------------------------------- runner


def run_logged(cmd: Sequence[str], log_path: Path, timeout_s: float = 0.0, echo: bool = True,
               cwd: Optional[Path] = None, env: Optional[Mapping[str, str]] = None) -> int:
    """Run `cmd`, teeing its merged stdout/stderr into `log_path` (and to our
    stdout when `echo`).  The time limit (0: none) is enforced while the child
    runs: on expiry its whole process group gets SIGTERM, then SIGKILL after
    5 s, and the result is 124.  Ctrl-C stops the group and re-raises."""
    log_path = Path(log_path)
    log_path.parent.mkdir(parents=True, exist_ok=True)
    with log_path.open("w", encoding="utf-8", errors="replace") as log:
        child = subprocess.Popen(list(cmd), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                                 errors="replace", bufsize=1, start_new_session=True, cwd=cwd,
                                 env=None if env is None else dict(env))

        def pump() -> None:
            assert child.stdout is not None
            for line in child.stdout:
                log.write(line)
                log.flush()
                if echo:
                    sys.stdout.write(line)
                    sys.stdout.flush()

        reader = threading.Thread(target=pump, daemon=True)
        reader.start()
        try:
            rc = child.wait(timeout=timeout_s if timeout_s > 0 else None)
        except subprocess.TimeoutExpired:
            _stop_group(child)
            rc = 124
        except KeyboardInterrupt:
            _stop_group(child)
            raise
        finally:
            reader.join(timeout=10)
    return rc


def _stop_group(child: subprocess.Popen) -> None:
    for sig, grace in ((signal.SIGTERM, 5.0), (signal.SIGKILL, 5.0)):
        try:
            os.killpg(child.pid, sig)
        except ProcessLookupError:
            return
        try:
            child.wait(timeout=grace)
            return
        except subprocess.TimeoutExpired:
            continue
