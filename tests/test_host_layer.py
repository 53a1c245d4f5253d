"""The hemo::HemoCell host layer on the CPU: its XML reader returns what a
standard XML parser reads from the reference's own configuration and material
files (comments, declarations and whitespace included), a missing element is
reported as missing (std::invalid_argument in the reference, config.cpp:85),
and the drivers print their usage without touching a device."""
import os
import subprocess
import xml.etree.ElementTree as ET

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hemocell_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden", "stretch_cell")


def _read(path, keys):
    r = subprocess.run([os.path.join(BIN, "config_read"), path] + keys, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


@pytest.mark.parametrize("name", ["config_stretch_cell.xml", "stretch_RBC.xml"])
def test_xml_reader_matches_etree(name):
    path = os.path.join(GOLDEN, name)
    root = ET.parse(path).getroot()
    keys, want = [], []
    for sec in root:
        for el in sec:
            if len(el) == 0:
                keys.append("%s/%s" % (sec.tag, el.tag))
                want.append((el.text or "").strip())
    assert len(keys) >= 5
    assert _read(path, keys) == want


def test_xml_reader_reports_missing_elements():
    path = os.path.join(GOLDEN, "config_stretch_cell.xml")
    assert _read(path, ["domain/dx", "domain/none", "nothing/dx"]) == ["0.5e-6", "<missing>", "<missing>"]


def test_material_numbers_parse(tmp_path):
    p = tmp_path / "m.xml"
    p.write_text('<?xml version="1.0" ?>\n<!-- c -->\n<hemocell><MaterialModel><kLink> 15.0 </kLink>'
                 '<!-- inline --><radius>3.91e-6</radius><InnerEdges><Edge> 1 2 </Edge><Edge>3 4</Edge>'
                 '</InnerEdges><empty/></MaterialModel></hemocell>\n')
    assert _read(str(p), ["MaterialModel/kLink", "MaterialModel/radius", "MaterialModel/InnerEdges/Edge",
                          "MaterialModel/empty"]) == ["15.0", "3.91e-6", "1 2", ""]


def _tree(el):
    return (el.tag, (el.text or "").strip(), [_tree(c) for c in el])


@pytest.mark.parametrize("name", ["config_stretch_cell.xml", "stretch_RBC.xml"])
def test_xml_writer_round_trips(name, tmp_path):
    """The <hemocell> element as saveCheckPoint() writes it into checkpoint.xml
    (core/hemoCellFields.cpp:270-276) parses back to the same tree, and a
    <Checkpoint> file routes cfg[...] through its <hemocell> (config/config.cpp:52-75)."""
    path = os.path.join(GOLDEN, name)
    r = subprocess.run([os.path.join(BIN, "config_read"), path, "--xml"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert _tree(ET.fromstring(r.stdout)) == _tree(ET.parse(path).getroot())
    chk = tmp_path / "checkpoint.xml"
    chk.write_text('<?xml version="1.0" ?>\n<Checkpoint>\n%s<General><Iteration>40</Iteration>'
                   "<OutDirectory>./tmp/</OutDirectory></General>\n</Checkpoint>\n" % r.stdout)
    root = ET.parse(path).getroot()
    keys = ["%s/%s" % (sec.tag, el.tag) for sec in root for el in sec if len(el) == 0]
    assert _read(str(chk), keys) == _read(path, keys)


@pytest.mark.parametrize("driver", ["stretch_cell", "load_particles", "checkpoint_run", "config_read"])
def test_drivers_print_usage(driver):
    r = subprocess.run([os.path.join(BIN, driver)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
