"""XML definition of a force (serialization/src/GluedForceProxy.cpp,
format version 1): round trips, the proxy's node layout read from text
written by hand in that layout, and the version check (CPU)."""
import math

import numpy as np
import pytest

from gluedmd_amd.glued import Force, GluedForce


def _force():
    f = Force(pbc=True, temperature=310.0)
    f.setForceGroup(3)
    d = f.add_distance([0, 1])
    t = f.add_dihedral([2, 3, 4, 5])
    c = f.add_coordination(list(range(6, 10)), list(range(10, 20)), r0=0.4, cutoff=1.2)
    p = f.add_path(list(range(4)), [np.random.default_rng(0).uniform(0, 1, 12).tolist() for _ in range(3)], 50.0)
    f.add_metad([d, t], [0.1, 0.35], 1.2, 10, grid_min=[0.0, -math.pi], grid_max=[2.0, math.pi], bins=[40, 36],
                periodic=[False, True], gamma=6.0)
    f.add_pbmetad([c, p], [0.2, 0.5], 1.0, 5, grid_min=[0.0, 1.0], grid_max=[10.0, 3.0], bins=[50, 20], gamma=8.0)
    f.add_harmonic([t], 10.0, 1.0 / 3.0, periodic=True)
    f.add_upper_wall(d, 1.5, 100.0, n=4, eps=0.1)
    f.add_opes([d, t], [0.05, 0.2], gamma=12.0, pace=100)
    return f


def test_xml_round_trip():
    f = _force()
    g = Force.deserialize_xml(f.serialize_xml())
    assert g.cvs == f.cvs and g.biases == f.biases
    assert g.getTemperature() == 310.0 and g.usesPeriodicBoundaryConditions() and g.getForceGroup() == 3
    assert g.getNumCollectiveVariables() == f.getNumCollectiveVariables()
    # parameters keep every bit (1/3, pi, path frames)
    assert g.biases[-3][2][1] == 1.0 / 3.0 and g.biases[0][2][-1] == math.pi


def test_xml_reads_the_proxy_layout():
    # GluedForceProxy.cpp:16-78: version, forceGroup, temperature, usesPBC;
    # CV {type, numAtoms, atom<j>, numParams, param<j>}; Bias {type, numCVs,
    # cv<j>, numParams, param<j>, numIntParams, intParam<j>}
    text = ('<Force type="GluedForce" version="1" forceGroup="0" temperature="300" usesPBC="0">'
            '<CollectiveVariables>'
            '<CV type="1" numAtoms="2" atom0="0" atom1="21" numParams="0"/>'
            '<CV type="3" numAtoms="4" atom0="4" atom1="6" atom2="8" atom3="14" numParams="0"/>'
            '</CollectiveVariables>'
            '<Biases>'
            '<Bias type="1" numCVs="1" cv0="0" numParams="2" param0="100" param1="0.5" numIntParams="0"/>'
            '<Bias type="3" numCVs="1" cv0="1" numParams="6" param0="1.2" param1="0.35" param2="6" '
            'param3="2.494" param4="-3.14159" param5="3.14159" numIntParams="3" intParam0="500" intParam1="100" '
            'intParam2="1"/>'
            '</Biases></Force>')
    f = Force.deserialize_xml(text)
    assert f.cvs == [(1, [0, 21], []), (3, [4, 6, 8, 14], [])]
    assert f.biases == [(1, [0], [100.0, 0.5], []),
                        (3, [1], [1.2, 0.35, 6.0, 2.494, -3.14159, 3.14159], [500, 100, 1])]
    assert f.getTemperature() == 300.0 and not f.usesPeriodicBoundaryConditions()
    assert isinstance(GluedForce.deserialize_xml(text), GluedForce)


def test_xml_version_and_unsupported_cvs():
    with pytest.raises(ValueError):
        Force.deserialize_xml('<Force type="GluedForce" version="2"><CollectiveVariables/><Biases/></Force>')
    with pytest.raises(ValueError):
        Force.deserialize_xml('<Force type="GluedForce" version="1"><CollectiveVariables>'
                              '<CV type="24" numAtoms="0" numParams="0" expression="x"/></CollectiveVariables>'
                              '<Biases/></Force>')


def test_xml_definition_evaluates_the_same(oracle):
    f = _force()
    g = Force.deserialize_xml(f.serialize_xml())
    pos = np.random.default_rng(5).uniform(0, 2, size=(24, 3)).astype(np.float32)
    ra = oracle.OracleContext(f, 24)
    rb = oracle.OracleContext(g, 24)
    ra.set_box(2.5)
    rb.set_box(2.5)
    a, b = ra.execute(pos), rb.execute(pos)
    assert np.array_equal(a["cv"], b["cv"]) and np.array_equal(a["bias_energies"], b["bias_energies"])
