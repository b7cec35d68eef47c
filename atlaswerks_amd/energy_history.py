"""EnergyHistory XML files, as the reference writes them (``*Energy.xml``).

``EnergyHistory::SaveXML`` (Algorithms/Common/EnergyHistory.cxx:295-316) builds a
TinyXML tree and prints it with ``TiXmlDocument::SaveFile``: one element per line,
four spaces of indentation per level, attributes in insertion order as
``name="value"``, childless elements as ``<name ... />``, numbers with ``%f`` /
``%d``.  Events (EnergyHistory.cxx:18-150):

* ``IterationEvent`` -> ``<Iteration Scale="s" Iter="i">`` holding the energy;
* ``StepSizeEvent`` -> ``<StepSizeChange val="..." />``;
* ``EnergyIncreaseEvent`` -> ``<EnergyIncrease />``;
* ``ReparameterizeEvent`` -> ``<Reparameterize>`` holding the energy.

Energies: ``Energy::ToXML`` (Energy.cxx:60-67) is ``<Energy val="total" />``;
``LDMMEnergy::ToXML`` (LDMMEnergy.cxx:212-238) adds ``<ImageEnergy val=... />``
and ``<VecEnergy val=... NTimeSteps=... />``, and its per-time-step loop sets
``val`` and ``t`` on the VecEnergy element again for every t (the VecStepEnergy
elements it creates are never attached), so the file carries the last time
step's energy in ``val`` and ``t = T - 1``.  Given the per-step energies, this
module writes the same; given only the mean vector energy, ``val`` is that mean.
"""


def _fmt(v):
    return "%f" % float(v)


class _Elem:
    def __init__(self, name, **attrs):
        self.name = name
        self.attrs = []  # insertion order; setting again replaces in place
        self.children = []
        for k, v in attrs.items():
            self.set(k, v)

    def set(self, key, value):
        for i, (k, _) in enumerate(self.attrs):
            if k == key:
                self.attrs[i] = (key, value)
                return
        self.attrs.append((key, value))

    def lines(self, depth, out):
        ind = "    " * depth
        head = ind + "<" + self.name + "".join(' %s="%s"' % kv for kv in self.attrs)
        if not self.children:
            out.append(head + " />")
            return
        out.append(head + ">")
        for c in self.children:
            c.lines(depth + 1, out)
        out.append(ind + "</" + self.name + ">")


class Energy:
    """Greedy / base energy (Energy.cxx): the total only."""

    def __init__(self, total):
        self.total = float(total)

    def to_elem(self):
        return _Elem("Energy", val=_fmt(self.total))


class LDMMEnergy(Energy):
    """LDMM energy (LDMMEnergy.cxx): image + vector energy over n_time_steps.

    vec_steps: the per-time-step vector energies (optional); vec: their mean
    (VecEnergy(), LDMMEnergy.cxx:57-64) when the steps are not available."""

    def __init__(self, image, vec=None, n_time_steps=0, vec_steps=None):
        if vec_steps is not None:
            vec_steps = [float(v) for v in vec_steps]
            n_time_steps = len(vec_steps)
            vec = sum(vec_steps) / max(n_time_steps, 1)
        self.image = float(image)
        self.vec = float(vec or 0.0)
        self.n_time_steps = int(n_time_steps)
        self.vec_steps = vec_steps
        super().__init__(self.image + self.vec)

    def to_elem(self):
        e = super().to_elem()
        e.children.append(_Elem("ImageEnergy", val=_fmt(self.image)))
        v = _Elem("VecEnergy", val=_fmt(self.vec), NTimeSteps="%d" % self.n_time_steps)
        if self.vec_steps is not None:
            for t, vs in enumerate(self.vec_steps):
                v.set("val", _fmt(vs))
                v.set("t", "%d" % t)
        e.children.append(v)
        return e


class EnergyHistory:
    """The event list of one registration or atlas (EnergyHistory.h)."""

    def __init__(self):
        self.events = []

    def add_iteration(self, scale, it, energy):
        """energy: a float (greedy), an Energy, or an LDMMEnergy."""
        if not isinstance(energy, Energy):
            energy = Energy(energy)
        self.events.append(("iter", int(scale), int(it), energy))

    def add_step_size(self, step):
        self.events.append(("step", float(step)))

    def add_energy_increase(self):
        self.events.append(("increase",))

    def add_reparameterize(self, energy):
        self.events.append(("reparam", energy))

    def to_xml(self):
        root = _Elem("EnergyHistory")
        for ev in self.events:
            if ev[0] == "iter":
                el = _Elem("Iteration", Scale="%d" % ev[1], Iter="%d" % ev[2])
                el.children.append(ev[3].to_elem())
            elif ev[0] == "step":
                el = _Elem("StepSizeChange", val=_fmt(ev[1]))
            elif ev[0] == "increase":
                el = _Elem("EnergyIncrease")
            else:
                el = _Elem("Reparameterize")
                el.children.append(ev[1].to_elem())
            root.children.append(el)
        out = []
        root.lines(0, out)
        return "\n".join(out) + "\n"

    def save_xml(self, path):
        with open(path, "w") as f:
            f.write(self.to_xml())

    @classmethod
    def from_levels(cls, levels):
        """levels: [(scale index, [energy per iteration])] -> one IterationEvent per
        iteration, as GreedyAtlasBuilder::FinishIteration (GreedyAtlasBuilder.txx:340-348)
        and DeformationIteratorInterface (DeformationIteratorInterface.cxx:74-82) add."""
        h = cls()
        for scale, energies in levels:
            for i, e in enumerate(energies):
                h.add_iteration(scale, i, e)
        return h
