"""User parameters of an experiment (interface of core/parameters.py).

``UserParameters()`` loads ``defaults.json``; each category is a dictionary
that accepts new values for existing keys only and is frozen once ``Nyles``
is created.  ``check()`` validates types and allowed values with the rules
of the reference (core/parameters.py:149-247), including the power-of-two
grid sizes.  The ``device`` category holds the settings of the GPU time
step (precision mode and the multigrid stop test, whose reference values
are hard-wired in the mgfor library).
"""
import datetime
import json
import os

CATEGORIES = ["model", "physics", "IO", "animation", "time", "discretization", "MPI",
              "multigrid", "device"]
ATTRIBUTES = ["type", "default", "avail", "doc"]


class UserParameterError(Exception):
    """An error occured with the user-set parameters."""


class DefaultsFileError(Exception):
    def __init__(self, defaults_file, arg):
        self.args = ["bad format of file {}: {}".format(defaults_file, arg)]


class InextensibleDict(dict):
    """A dict whose keys are fixed; all instances freeze together."""

    frozen = False

    @classmethod
    def freeze(cls):
        cls.frozen = True

    def __setitem__(self, key, item):
        if self.frozen:
            raise UserParameterError(
                "not possible to modify parameters after the creation of Nyles.")
        if key not in self:
            raise UserParameterError(
                "not possible to add new key {!r} to the parameters.".format(key))
        dict.__setitem__(self, key, item)


class UserParameters(object):
    DEFAULTS_FILE = "defaults.json"
    TYPES = {
        "str": str,
        "int": int,
        "float": (float, int),
        "bool": bool,
        "dict": dict,
        "list or string": (list, tuple, str),
    }

    def __init__(self):
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), self.DEFAULTS_FILE)
        with open(path) as f:
            defaults = json.load(f)
        self.check_defaults(defaults)
        for cat in CATEGORIES:
            setattr(self, cat, InextensibleDict({
                name: attrs["default"] for name, attrs in defaults[cat].items()}))
        self.documentations, self.options, self.types = {}, {}, {}
        for params in defaults.values():
            for name, attrs in params.items():
                self.documentations[name] = attrs["doc"]
                self.options[name] = attrs["avail"]
                self.types[name] = attrs["type"]

    def help(self, parameter):
        if parameter in self.documentations:
            return self.documentations[parameter]
        raise ValueError("invalid parameter: {!r}".format(parameter))

    def possible_values(self, parameter):
        if parameter in self.options:
            return self.options[parameter]
        raise ValueError("invalid parameter: {!r}".format(parameter))

    def view_parameters(self):
        out = {}
        for cat in CATEGORIES:
            out.update(getattr(self, cat))
        return out

    def check(self):
        powers = [2 ** n for n in range(datetime.datetime.now().year - 2000)]
        for name, value in self.view_parameters().items():
            ptype = self.types[name]
            if not isinstance(value, self.TYPES[ptype]):
                raise UserParameterError("parameter {} must be {} {}, not {}".format(
                    name, "an" if ptype == "int" else "a", ptype, type(value)))
            opts = self.options[name]
            if isinstance(opts, list):
                if value not in opts:
                    raise UserParameterError("parameter {} must be one of {}, not {!r}".format(
                        name, opts, value))
            elif opts == "any":
                pass
            elif opts == "> 0.0":
                if not value > 0.0:
                    raise UserParameterError("parameter {} must be positive".format(name))
            elif opts in (">= 0.0", ">= 0"):
                if not value >= 0.0:
                    raise UserParameterError("parameter {} must be non-negative".format(name))
            elif opts == ">= 1":
                if not value >= 1:
                    raise UserParameterError("parameter {} must be at least 1".format(name))
            elif opts in ("2^n", "[3 *] 2^n"):
                ok = value in powers or (opts == "[3 *] 2^n" and value / 3 in powers)
                if not ok:
                    if value < max(powers):
                        raise UserParameterError(
                            "parameter {} must be a power of 2{}".format(
                                name, "" if opts == "2^n" else " or 3 times a power of 2"))
                    raise UserParameterError(
                        "parameter {} is very large; if you are sure to use it, extend the "
                        "list of powers of 2".format(name))
            elif opts == "any valid filename":
                if "/" in value:
                    raise UserParameterError('parameter {} must not contain a "/"'.format(name))
            elif name == "variables_in_history":
                if not isinstance(value, (list, tuple)) and value not in (
                        "all", "prognostic", "p+p"):
                    raise UserParameterError(
                        "value {!r} of parameter {} not understood".format(value, name))
            elif name == "datadir":
                pass
            else:
                print("Warning: cannot check parameter", name)
        for x in "xyz":
            if self.MPI["np" + x] > self.discretization["global_n" + x]:
                raise UserParameterError(
                    "parameter np{} cannot be larger than global_n{}".format(x, x))

    @classmethod
    def check_defaults(cls, defaults):
        for cat in CATEGORIES:
            if cat not in defaults:
                raise DefaultsFileError(cls.DEFAULTS_FILE, "category {} missing".format(cat))
        seen = set()
        for cat, params in defaults.items():
            if cat not in CATEGORIES:
                raise DefaultsFileError(cls.DEFAULTS_FILE, "unknown category {}".format(cat))
            for name, attrs in params.items():
                if name in seen:
                    raise DefaultsFileError(cls.DEFAULTS_FILE,
                                            "parameter {} appears twice".format(name))
                seen.add(name)
                for a in ATTRIBUTES:
                    if a not in attrs:
                        raise DefaultsFileError(
                            cls.DEFAULTS_FILE,
                            "attribute {} missing for parameter {}".format(a, name))
                for a in attrs:
                    if a not in ATTRIBUTES:
                        raise DefaultsFileError(
                            cls.DEFAULTS_FILE,
                            "unknown attribute {} for parameter {}".format(a, name))
                if attrs["type"] not in cls.TYPES:
                    raise DefaultsFileError(
                        cls.DEFAULTS_FILE,
                        "unknown type {} of parameter {}".format(attrs["type"], name))

    def freeze(self):
        InextensibleDict.freeze()
