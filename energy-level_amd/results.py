"""Versioned data descriptors of a result, as ``describe_result``
(tomisnot/energy-level src/core/manybody/results.py:11-137) attaches them.

``data_schema = {version: 1, quantities: {path: {quantity, axes, unit, ...}},
axes: {...}, approximation}`` names what each array of a result is, along
which axes and in which unit, so that a consumer plots by quantity instead of
by solver; ``sites`` lists the calib's atoms.  The quantity names, axes and
units are the reference's table (the result contract); the descriptors refer
to paths in the result and copy no numerical buffer.
"""
from __future__ import annotations

import numpy as np

__all__ = ['describe_result']


def _at(result, path):
    node = result
    for key in path.split('.'):
        node = node.get(key) if isinstance(node, dict) else None
    return node


def describe_result(result, calib=None, *, loss_kind='norm_drift', approximation=None):
    """Adds ``data_schema`` (and ``sites`` with a calib) to ``result`` in place
    and returns it."""
    out = result
    n = len(calib['atoms']) if calib else int(out.get('n_atoms', out.get('n_sites', 0)) or 0)
    quantities = {}

    def add(path, quantity, axes, unit='1', **extra):
        value = _at(out, path)
        if value is None:
            return
        arr = np.asarray(value)
        if arr.size == 0:
            return
        quantities[path] = dict(quantity=quantity, axes=list(axes), unit=unit, **extra)
        if (path in ('S', 'gap', 'observables.phase') and np.issubdtype(arr.dtype, np.number)
                and not np.isfinite(arr).all()):
            quantities[path]['missing_values'] = 'nonfinite'

    if out.get('pops') is not None:
        p = np.asarray(out['pops'])
        labels = out.get('states') or []
        if p.ndim != 2 or p.shape[1] != len(labels):
            raise ValueError('result contract: every column of pops needs its states entry')
        kind = 'basis' if labels and 'index' in labels[0] else 'site'
        out['pops_kind'] = 'basis_probability' if kind == 'basis' else 'site_occupation'
        add('pops', out['pops_kind'], ('time', kind), summary=dict(mode='top', n=8, labels='states'))
    for key in ('pops_atom', 'occupations'):
        add(key, 'site_occupation', ('time', 'site'))
    add('basis_probabilities', 'basis_probability', ('time', 'basis'))
    add('stderr', 'standard_error', ('time', 'site'), of='pops_atom')
    add('S', 'entanglement_entropy', ('time',) if out.get('t') is not None else (), 'nat',
        **({'summary': dict(mode='scalar')} if out.get('t') is None else {}))
    corr_3d = (out.get('betas') is not None and out.get('corr') is not None
               and np.asarray(out['corr']).ndim == 3)
    add('corr', 'connected_occupation_correlation',
        ('beta', 'site', 'site') if corr_3d else ('site', 'site'),
        **({'summary': dict(mode='first_last_axis', axis=0)} if corr_3d else {}))
    add('energy', 'energy', ('time',), 'MHz')
    add('E0', 'ground_state_energy', (), 'MHz', summary=dict(mode='scalar'))
    for key, quantity, unit in (('E_per_site', 'energy_per_site', 'MHz'),
                                ('E_per_cell', 'energy_per_cell', 'MHz'),
                                ('S_per_bond', 'entanglement_entropy_per_bond', 'nat')):
        add(key, quantity, (), unit)
    add('entanglement_entropy', 'entanglement_entropy', ('site',), 'nat')
    add('corr_length', 'correlation_length', (), 'site')
    add('chi', 'mps_bond_dimension', (), '1', summary=dict(mode='scalar'))
    for key, quantity, unit in (('trunc_err', 'truncation_error', '1'),
                                ('unit_cell_size', 'unit_cell_size', '1'),
                                ('energy_expectation', 'energy_expectation', 'MHz'),
                                ('energy_variance', 'energy_residual_variance', 'MHz^2'),
                                ('permutation_overlap', 'site_permutation_expectation', '1'),
                                ('projection_weight', 'symmetry_projection_weight', '1')):
        add(key, quantity, (), unit)
    add('gap', 'energy_gap', (), 'MHz', summary=dict(mode='scalar'))
    add('evals', 'eigenenergy', ('eigenstate',), 'MHz')
    for key, quantity, unit in (('E', 'energy', 'MHz'), ('C', 'heat_capacity', '1'),
                                ('F', 'free_energy', 'MHz')):
        add(key, quantity, ('beta',), unit, summary=dict(mode='first_last'))
    add('E_err', 'standard_error', ('beta',) if out.get('betas') is not None else (), 'MHz',
        of='E')
    add('rho', 'density_of_states', ('energy',), '1/MHz')
    # source spectrum: the spectral function on the `energies` grid; the
    # quadrature weights live on the Ritz nodes (another axis) and stay
    # undeclared, as the reference leaves them
    add('spectral', 'spectral_function', ('energy',), '1/MHz')
    add('source_norm', 'source_state_weight', (), '1')
    for key, quantity, unit in (('gaps', 'energy_gap', 'MHz'), ('ns', 'occupation_density', '1'),
                                ('Om_t', 'drive_amplitude', 'MHz'), ('de_t', 'detuning', 'MHz')):
        add(key, quantity, ('tau',), unit)
    if out.get('t') is not None:
        n_atom_summary = {}
    elif out.get('betas') is not None:
        n_atom_summary = {'summary': dict(mode='rows', n=8)}
    else:
        n_atom_summary = {'summary': dict(mode='list')}
    add('n_atom', 'site_occupation', ('time', 'site') if out.get('t') is not None else
        ('beta', 'site') if out.get('betas') is not None else ('site',), **n_atom_summary)
    for key in ('sx', 'sy', 'sz', 'sp', 'sm', 'phase'):
        add('observables.' + key, key, ('time', 'site'), 'rad' if key == 'phase' else '1')
    for key, axes, unit in (('xy_corr', ('site', 'site'), '1'),
                            ('link_corr', ('time', 'link'), '1'),
                            ('loop_phase_series', ('time', 'loop'), 'rad'),
                            ('loop_winding_series', ('time', 'loop'), '1')):
        add('observables.' + key, key, axes, unit)
    timed = out.get('t') is not None
    add('observables.pauli_strings', 'pauli_string_expectation',
        ('time', 'string') if timed else ('string',), '1',
        summary=dict(mode='final' if timed else 'head', n=12,
                     labels=('observables.pauli_string_labels' if timed else 'pauli_string_labels')),
        **({'items': dict(labels='observables.pauli_string_labels')} if timed else {}))
    add('pauli_strings', 'pauli_string_expectation', ('string',), '1',
        summary=dict(mode='head', n=12, labels='pauli_string_labels'))
    add('observables.loop_phase', 'loop_phase', ('loop',), 'rad')
    add('observables.loop_winding', 'loop_winding', ('loop',))
    if out.get('loss') is not None:
        out[loss_kind] = out['loss']
        add(loss_kind, loss_kind, ('time',))
        add('loss', loss_kind, ('time',), alias_of=loss_kind)
    add('survival_weight', 'no_jump_survival_weight', ('time',))
    add('jump_counts', 'jump_count', ('trajectory',))
    add('spins', 'classical_spin', ('time', 'site', 'component'))
    out['data_schema'] = dict(
        version=1, quantities=quantities,
        axes=dict(time=dict(field='t', unit='ns'), beta=dict(field='betas', unit='1/MHz'),
                  energy=dict(field='energies', unit='MHz'), tau=dict(field='taus', unit='1'),
                  site=dict(field='sites', size=n), link=dict(field='observables.link_pairs'),
                  loop=dict(field='loops'), trajectory=dict(field='jump_counts'),
                  eigenstate=dict(field='evals'), component=dict(values=['x', 'y', 'z']),
                  basis=dict(field=('basis_states' if out.get('basis_states') is not None
                                    else 'states'), bit_order='site0-msb')),
        approximation=approximation)
    if calib:
        out['sites'] = [dict(site_id=a.get('id', i), **({'pos': list(a['pos'])} if 'pos' in a else {}))
                        for i, a in enumerate(calib['atoms'])]
    return out
