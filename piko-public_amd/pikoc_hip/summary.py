"""Pipe summary and kernel plan of pikoc_hip.

The summary holds what pikoc's PipeSummary / stageSummary hold
(pikoc/include/PikoSummary.hpp:140-357): per stage its bin size, threads per
tile, primitive types, assignBin / schedule / wait policies, ports and
neighbours; per pipe the stages, the state and input types and the drain
stages.  The kernel plan splits the stages into launches with the fusion
legality of PipeSummary::canFuse (pikoc/src/PikoSummary.cpp:686-755) and the
kernel list of makeKernelList (pikoc/src/pikoc.cpp:32-67): a kernel 0 that
runs the first stage's assignBin over the input, then one process kernel per
stage (preceded by a schedule kernel when the schedule phase is not a bare
policy), a stage fusing into its producer's kernel when canFuse allows.
"""
import copy


class StageSummary:
    def __init__(self, name, type):
        self.name, self.type = name, type
        self.binsize = (0, 0)          # (0, 0): one full-screen bin
        self.threads_per_tile = 1
        self.prim_in = self.prim_out = ""
        self.assign_policy = "Empty"   # Custom | InBin | BoundingBox | Position | Empty
        self.assign_trivial = False
        self.sched_policy = "Custom"   # LoadBalance | DirectMap | Serialize | All | Custom
        self.sched_trivial = False
        self.wait_policy = "None"      # None | Batch | EndStage | EndBin | Custom
        self.tile_split_size = 0
        self.max_out_prims = 0
        self.process_policy = "Empty"
        # normalised source of the class body and of its assignBin /
        # schedule / process bodies (None where a phase is absent)
        self.source = {}
        # EndStage wait: the stage whose end this stage waits for (pikoc's
        # schedule endStageName / endStagePtr, PikoSummary.hpp:198-199, set
        # from the name in PipeSummary's link pass, PikoSummary.cpp:609-611)
        self.end_stage_name = ""
        self.end_stage = None
        self.next_by_port = {}         # out port -> [StageSummary]
        self.prev = []
        self.next = []
        self.loop = False
        self.dist_from_drain = None

    def copy_as(self, name):
        s = copy.copy(self)
        s.name = name
        s.next_by_port, s.prev, s.next = {}, [], []
        return s

    def describe(self):
        return ("%s (%s): bins %dx%d, %d threads/tile, %s -> %s, assignBin %s, schedule %s%s, wait %s"
                % (self.name, self.type, self.binsize[0], self.binsize[1], self.threads_per_tile, self.prim_in,
                   self.prim_out, self.assign_policy, self.sched_policy,
                   "(%d)" % self.tile_split_size if self.tile_split_size else "", self.wait_policy))


class PipeSummary:
    def __init__(self, name, filename):
        self.name, self.filename = name, filename
        self.stages = []
        self.const_state_type = self.mutable_state_type = self.input_type = ""
        self.input_fields = None
        self.macros = {}
        self.drain = []
        self.order = []
        self.has_loop = False

    def stage(self, name):
        for s in self.stages:
            if s.name == name:
                return s
        raise KeyError(name)

    def add_stage(self, s):
        self.stages.append(s)

    def connect(self, out_name, in_name, out_port, in_port):
        a, b = self.stage(out_name), self.stage(in_name)
        if not (0 <= out_port < 5 and 0 <= in_port < 5):
            raise ValueError("ports are 0..4")
        a.next_by_port.setdefault(out_port, []).append(b)
        if b not in a.next:
            a.next.append(b)
        if a not in b.prev:
            b.prev.append(a)

    def process_links(self):
        """Loops, drain stages and their distances, and the stage order
        (sources first, then breadth-first along the ports)."""
        for s in self.stages:
            s.loop = s in s.next
            # PikoSummary.cpp:609-611: an EndStage wait names the stage it waits for
            s.end_stage = None
            if s.wait_policy == "EndStage" and s.end_stage_name:
                s.end_stage = self.stage(s.end_stage_name)
        self.has_loop = any(s.loop for s in self.stages)
        self.drain = [s for s in self.stages if not [n for n in s.next if n is not s]]
        for s in self.drain:
            self._drain_dist(s, 0)
        sources = [s for s in self.stages if not [p for p in s.prev if p is not s]]
        order, queue = [], list(sources)
        while queue:
            s = queue.pop(0)
            if s in order:
                continue
            order.append(s)
            for port in sorted(s.next_by_port):
                for n in s.next_by_port[port]:
                    if n is not s and n not in order:
                        queue.append(n)
        order += [s for s in self.stages if s not in order]
        self.order = order

    def _drain_dist(self, s, d):
        if s.dist_from_drain is not None and s.dist_from_drain <= d:
            return
        s.dist_from_drain = d
        for p in s.prev:
            if p is not s:
                self._drain_dist(p, d + 1)

    @staticmethod
    def can_fuse(s1, s2, done):
        """PipeSummary::canFuse (PikoSummary.cpp:686-755) for the one
        schedule this backend plans.  `done`: the stages whose kernels have
        been planned before s1's."""
        if s1.type == s2.type:
            return False
        if len([p for p in s2.prev if p is not s2]) > 1:
            return False
        if len(s1.next) > 1:
            return False
        # an unresolved EndStage dependency blocks fusion (PikoSummary.cpp:707-709)
        if s2.end_stage is not None and not any(d is s2.end_stage for d in done):
            return False
        same_core = ((s2.sched_policy == s1.sched_policy and s2.assign_policy == "InBin"
                      and s1.sched_policy == "DirectMap")
                     or (s2.sched_policy == s1.sched_policy and s1.sched_policy == "Serialize")
                     or (s2.sched_policy == s1.sched_policy and s2.assign_policy == "InBin"
                         and s1.sched_policy == "LoadBalance"))
        if s2.binsize == s1.binsize and same_core and s1.tile_split_size == s2.tile_split_size:
            return True
        return (s1.sched_policy == "All" and s2.sched_policy == "All" and s2.assign_policy == "InBin"
                and s1.binsize == s2.binsize and s2.wait_policy == "None"
                and s1.tile_split_size == s2.tile_split_size)

    def kernel_plan(self):
        """[(kernel name, [stage phases])] in launch order."""
        plan = [("kernel0", ["%s.assignBin(input)" % self.order[0].name])]
        k = 1
        groups = []
        for s in self.order:
            if groups and self.can_fuse(groups[-1][-1], s, [d for g in groups[:-1] for d in g]):
                groups[-1].append(s)
            else:
                groups.append([s])
        for g in groups:
            head = g[0]
            if not head.sched_trivial:
                plan.append(("kernel%d" % k, ["%s.schedule" % head.name]))
                k += 1
            phases = []
            for s in g:
                phases.append("%s.process%s" % (s.name, " (loops until empty)" if s.loop else ""))
                for port in sorted(s.next_by_port):
                    for n in s.next_by_port[port]:
                        phases.append("emit[%d] -> %s.assignBin" % (port, n.name))
            plan.append(("kernel%d" % k, phases))
            k += 1
        return plan

    def display(self):
        lines = ["%s (%s)" % (self.name, self.filename), "=" * (len(self.name) + len(self.filename) + 3),
                 "* Drain stages: " + " ".join("[%s]" % s.name for s in self.drain),
                 "* Stages:"]
        for s in self.order:
            lines.append("  " + s.describe())
            for port in sorted(s.next_by_port):
                lines.append("    out port %d -> %s" % (port, ", ".join(n.name for n in s.next_by_port[port])))
        lines.append("* Kernel plan:")
        for name, phases in self.kernel_plan():
            lines.append("  %s: %s" % (name, "; ".join(phases)))
        return "\n".join(lines)
