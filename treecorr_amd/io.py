"""Output files of the correlation classes: TreeCorr's ASCII format.

Restates treecorr/writer.py AsciiWriter (:23-97), treecorr/reader.py AsciiReader
(:28-200), treecorr/util.py parse_file_type / make_writer / make_reader (:109-165) and
treecorr/config.py make_minimal_config (:423-433).  A file is a sequence of groups; each
group is an optional '## name' line, an optional '## {params}' line, a header line
'#  col1  col2 ...' (names centred in columns of precision + 8 characters) and the rows
in '%{precision+8}.{precision}e'.  FITS and HDF5 files need fitsio / h5py, which this
engine does not use.

The params line is a Python dict literal.  It is read with a literal parser that
accepts numbers, strings, tuples, lists, dicts, booleans, None and numpy scalar
spellings such as np.float64(1.5) (as numpy 2 prints them); it never evaluates code.
"""
import ast
import os

import numpy as np


def parse_file_type(file_type, file_name, output=False):
    """treecorr/util.py:109-133"""
    if file_type is None:
        ext = os.path.splitext(file_name)[1].lower()
        if ext.startswith('.fit'):
            file_type = 'FITS'
        elif ext.startswith('.hdf') or ext.startswith('.h5'):
            file_type = 'HDF'
        elif not output and ext.startswith('.par'):
            file_type = 'Parquet'
        else:
            file_type = 'ASCII'
    return file_type.upper()


def _check_ascii(file_type):
    if file_type in ('FITS', 'HDF', 'PARQUET'):
        raise NotImplementedError("%s files need a library that treecorr_amd does not use; "
                                  "write ASCII files instead" % file_type)
    if file_type != 'ASCII':
        raise ValueError("Invalid file_type %s" % file_type)


def make_writer(file_name, precision=4, file_type=None):
    _check_ascii(parse_file_type(file_type, file_name, output=True))
    return AsciiWriter(file_name, precision=precision)


def make_reader(file_name, file_type=None):
    _check_ascii(parse_file_type(file_type, file_name))
    return AsciiReader(file_name)


def make_minimal_config(config, valid_params):
    """The config without the keys that are not valid or hold their default value."""
    return {k: v for k, v in config.items() if k in valid_params and v != valid_params[k]}


_NP_SCALARS = {'float64': float, 'float32': float, 'int64': int, 'int32': int, 'bool_': bool,
               'str_': str, 'float_': float, 'int_': int}


def literal(text):
    """A Python literal, with numpy scalar calls (np.float64(x)) taken as their value."""
    def conv(node):
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and \
                isinstance(node.func.value, ast.Name) and node.func.value.id in ('np', 'numpy') \
                and node.func.attr in _NP_SCALARS and len(node.args) == 1 and not node.keywords:
            return _NP_SCALARS[node.func.attr](conv(node.args[0]))
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = conv(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.Tuple):
            return tuple(conv(e) for e in node.elts)
        if isinstance(node, ast.List):
            return [conv(e) for e in node.elts]
        if isinstance(node, ast.Dict):
            return {conv(k): conv(v) for k, v in zip(node.keys, node.values)}
        if isinstance(node, ast.Name) and node.id in ('inf', 'nan'):
            return float(node.id)
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and \
                node.func.id == 'float' and len(node.args) == 1:
            return float(conv(node.args[0]))
        raise ValueError("not a literal: %s" % ast.dump(node)[:80])
    return conv(ast.parse(text.strip(), mode='eval').body)


def _plain(v):
    """Python scalars for the params line (numpy scalars print as np.float64(...))."""
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, tuple):
        return tuple(_plain(x) for x in v)
    return v


class AsciiWriter:
    """treecorr/writer.py:23-97"""

    def __init__(self, file_name, *, precision=4):
        self.file_name = file_name
        self.set_precision(precision)
        self._file = None
        d = os.path.dirname(file_name)
        if d and not os.path.exists(d):
            os.makedirs(d)

    def set_precision(self, precision):
        self.precision = precision
        self.width = precision + 8
        self.fmt = '%%%d.%de' % (self.width, self.precision)

    @property
    def file(self):
        if self._file is None:
            raise RuntimeError('Illegal operation when not in a "with" context')
        return self._file

    def write(self, col_names, columns, *, params=None, ext=None):
        ncol = len(col_names)
        data = np.empty((len(columns[0]), ncol))
        for i, col in enumerate(columns):
            data[:, i] = col
        # the first name is one narrower to leave room for the '#'
        header = ("#" + "{:^%d}" % (self.width - 1) +
                  " {:^%d}" % (self.width) * (ncol - 1) + "\n").format(*col_names)
        if ext is not None:
            self.file.write(('## %s\n' % ext).encode())
        if params is not None:
            params = {k: _plain(v) for k, v in params.items()}
            self.file.write(('## %r\n' % (params,)).encode())
        self.file.write(header.encode())
        np.savetxt(self.file, data, fmt=self.fmt)

    def write_array(self, data, *, ext=None):
        if ext is not None:
            self.file.write(('## %s\n' % ext).encode())
        np.savetxt(self.file, data, fmt=self.fmt)

    def __enter__(self):
        self._file = open(self.file_name, 'wb')
        return self

    def __exit__(self, *exc):
        self._file.close()
        self._file = None


class AsciiReader:
    """treecorr/reader.py:28-200: the groups of a file, read in order."""

    def __init__(self, file_name):
        self.file_name = file_name
        self._file = None
        self.col_names = None

    @property
    def file(self):
        if self._file is None:
            raise RuntimeError('Illegal operation when not in a "with" context')
        return self._file

    def __contains__(self, ext):
        # ASCII files have no extensions: only None is valid
        return ext is None

    def read_params(self, *, ext=None):
        header = self.file.readline()
        params = {}
        if header[1] == '#':
            tokens = header[2:].split()
            if tokens[0][0] != '{':
                name = tokens[0]
                if ext is not None and name != ext:
                    raise OSError("Mismatch in group names.  Expected %s, found %s" % (ext, name))
                header = self.file.readline()
            if header[1] == '#':
                params = literal(header[2:])
                header = self.file.readline()
        self.col_names = header[1:].split()
        return params

    def read_data(self, *, ext=None, max_rows=None):
        rows = []
        while max_rows is None or len(rows) < max_rows:
            pos = self.file.tell()
            line = self.file.readline()
            if not line:
                break
            if line.startswith('#'):
                self.file.seek(pos)
                break
            rows.append([float(t) for t in line.split()])
        data = np.array(rows, dtype=float).reshape(len(rows), len(self.col_names))
        out = np.empty(len(rows), dtype=[(c, float) for c in self.col_names])
        for i, c in enumerate(self.col_names):
            out[c] = data[:, i]
        return out

    def read_array(self, shape, *, ext=None):
        line = self.file.readline()
        if line.startswith('##'):
            line = self.file.readline()
        rows = [[float(t) for t in line.split()]]
        for _ in range(shape[0] - 1):
            rows.append([float(t) for t in self.file.readline().split()])
        return np.array(rows, dtype=float).reshape(shape)

    def __enter__(self):
        self._file = open(self.file_name, 'r')
        return self

    def __exit__(self, *exc):
        self._file.close()
        self._file = None
