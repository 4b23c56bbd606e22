"""Regenerates tractor_amd/sersic_table.py from the reference's
tractor/sersic.py: the `fits` list of SersicMixture.__init__ (sersic.py:
40-221) and `self.cores` (223-246) are read as literals from the parsed
source (nothing is run) and written out as data.  A data recipe, kept with
the tests that check the table.  Usage:
    python tests/golden/gen_sersic_table.py [/root/reference/tractor/sersic.py]
"""
import ast
import os
import sys


def read_tables(path):
    tree = ast.parse(open(path).read())
    fits = cores = None
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign):
            t = node.targets[0]
            if isinstance(t, ast.Name) and t.id == 'fits':
                fits = ast.literal_eval(node.value)
            if isinstance(t, ast.Attribute) and t.attr == 'cores':
                cores = ast.literal_eval(node.value)
    if not fits or not cores:
        raise ValueError('no fits / cores tables in %s' % path)
    return fits, cores


def _tup(xs):
    return ', '.join(repr(x) for x in xs) + (',' if len(xs) == 1 else '')


def render(fits, cores):
    out = ['"""Sersic-profile mixture table (the fits of SersicMixture,',
           'tractor/sersic.py:40-221, and the core fractions, 223-246): per range',
           '[lo, hi) of the Sersic index, nodes (index, amplitudes, variances) with a',
           'fixed number of components; CORES: (index, 1 - softened / true central',
           'flux) for n >= 1.  Data, kept as the reference tabulates it."""', '', 'RANGES = (']
    for lo, hi, grid in fits:
        out.append('    (%r, %r, (' % (lo, hi))
        for s, a, v in grid:
            out.append('        (%r, (%s), (%s)),' % (s, _tup(a), _tup(v)))
        out.append('    )),')
    out += [')', '', 'CORES = (']
    for i in range(0, len(cores), 4):
        out.append('    ' + ' '.join('(%r, %r),' % c for c in cores[i:i + 4]))
    out.append(')')
    return '\n'.join(out) + '\n'


if __name__ == '__main__':
    src = sys.argv[1] if len(sys.argv) > 1 else '/root/reference/tractor/sersic.py'
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    dst = os.path.join(root, 'tractor_amd', 'sersic_table.py')
    text = render(*read_tables(src))
    with open(dst, 'w') as f:
        f.write(text)
    print('wrote', dst)
