"""INI configuration with the accessors of frfs/inifile.py:19-141.

Only the behaviour the DGFS hot path reads is kept: get/getfloat/getint/
getbool with defaults that are written back, items/items_as, hasopt, set.
"""
import os
from configparser import ConfigParser, NoOptionError, NoSectionError

_sentinel = object()


class Inifile:
    def __init__(self, inistr=None):
        self._cp = cp = ConfigParser(inline_comment_prefixes=[";"])
        cp.optionxform = str
        if inistr:
            cp.read_string(inistr)

    @staticmethod
    def load(file):
        if isinstance(file, str):
            with open(file) as fh:
                return Inifile(fh.read())
        return Inifile(file.read())

    def set(self, section, option, value):
        if not self._cp.has_section(section):
            self._cp.add_section(section)
        self._cp.set(section, option, str(value))

    def hasopt(self, section, option):
        return self._cp.has_option(section, option)

    def get(self, section, option, default=_sentinel):
        try:
            val = self._cp.get(section, option)
        except (NoSectionError, NoOptionError):
            if default is _sentinel:
                raise
            self.set(section, option, default)
            val = self._cp.get(section, option)
        return os.path.expandvars(val)

    def getexpr(self, section, option, default=_sentinel, subs={}):
        return self.get(section, option, default)

    def getfloat(self, section, option, default=_sentinel):
        return float(self.get(section, option, default))

    def getint(self, section, option, default=_sentinel):
        return int(self.get(section, option, default))

    def getbool(self, section, option, default=_sentinel):
        v = self.get(section, option, default)
        return str(v).strip().lower() in {"1", "true", "yes", "on"}

    def items(self, section):
        if not self._cp.has_section(section):
            return {}
        return dict(self._cp.items(section))

    def items_as(self, section, type):
        out = {}
        for k, v in self.items(section).items():
            try:
                out[k] = type(v)
            except ValueError:
                pass
        return out

    def tostr(self):
        import io
        buf = io.StringIO()
        self._cp.write(buf)
        return buf.getvalue()
