"""EnergyHistory XML writer (EnergyHistory.cxx:18-316, Energy.cxx:60-67,
LDMMEnergy.cxx:212-238): the text TinyXML prints for the reference's trees."""
import xml.etree.ElementTree as ET

from atlaswerks_amd.energy_history import EnergyHistory, LDMMEnergy


def test_greedy_history_text():
    h = EnergyHistory()
    h.add_iteration(0, 0, 4475.8377093504205)
    h.add_step_size(0.0125)
    h.add_iteration(0, 1, 3900.5)
    h.add_energy_increase()
    want = ("<EnergyHistory>\n"
            "    <Iteration Scale=\"0\" Iter=\"0\">\n"
            "        <Energy val=\"4475.837709\" />\n"
            "    </Iteration>\n"
            "    <StepSizeChange val=\"0.012500\" />\n"
            "    <Iteration Scale=\"0\" Iter=\"1\">\n"
            "        <Energy val=\"3900.500000\" />\n"
            "    </Iteration>\n"
            "    <EnergyIncrease />\n"
            "</EnergyHistory>\n")
    assert h.to_xml() == want
    ET.fromstring(h.to_xml())  # well formed


def test_ldmm_energy_attribute_overwrite():
    # LDMMEnergy::ToXML sets val and t on the VecEnergy element once per time step,
    # so the last step's energy and index remain
    h = EnergyHistory()
    h.add_iteration(1, 7, LDMMEnergy(10.0, vec_steps=[1.0, 2.0, 6.0]))
    root = ET.fromstring(h.to_xml())
    e = root.find("Iteration/Energy")
    assert e.get("val") == "13.000000"          # image + mean vector energy
    assert e.find("ImageEnergy").get("val") == "10.000000"
    v = e.find("VecEnergy")
    assert (v.get("val"), v.get("NTimeSteps"), v.get("t")) == ("6.000000", "3", "2")
    assert '<VecEnergy val="6.000000" NTimeSteps="3" t="2" />' in h.to_xml()


def test_from_levels_and_save(tmp_path):
    h = EnergyHistory.from_levels([(0, [3.0, 2.0]), (1, [1.5])])
    p = tmp_path / "GreedyEnergy.xml"
    h.save_xml(str(p))
    it = ET.parse(str(p)).getroot().findall("Iteration")
    assert [(i.get("Scale"), i.get("Iter")) for i in it] == [("0", "0"), ("0", "1"), ("1", "0")]
