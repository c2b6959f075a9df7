"""Loader for nova's on-disk (PO)MDP formats (mirrors python/nova/file_loader.py:31-587).

Two formats:
* raw CSV (load_raw_mdp, load_raw_pomdp): a header row of sizes, then S, T, (O,) R, (Z, B) rows
  in nova's layouts (python/nova/file_loader.py:74-213).
* Cassandra's .pomdp text format (load_cassandra), the subset nova reads
  (python/nova/file_loader.py:215-587): discount, states, actions, observations, start /
  start include / start exclude, and T, O, R entries given per action, per (action, state) row
  or per element, with '*' wildcards and the 'uniform' / 'identity' shorthands. Rewards are
  reduced to R(s, a): an entry sets R[s, a] to the mean of its value(s), later entries
  overriding earlier ones.

After a load the attributes hold numpy arrays in nova's layouts: S[n, m, ns] int32 (-1 pads a
row), T[n, m, ns], O[m, n, z], R[n, m], Z[r, rz] int32, B[r, rz] float32, plus n, ns, m, z, r,
rz, k, s0, ng, goals, gamma, horizon, Rmin, Rmax and epsilon = (Rmax - Rmin) / 1000.
"""
import csv
import math

import numpy as np


class FileLoader:
    def __init__(self):
        self.n = self.ns = self.m = self.z = self.r = self.rz = self.k = 0
        self.s0 = self.ng = 0
        self.goals = None
        self.gamma = 0.0
        self.horizon = 0
        self.Rmin = self.Rmax = 0.0
        self.S = self.T = self.O = self.R = self.Z = self.B = None
        self.epsilon = 0.0

    # ------------------------------------------------------------------ raw CSV formats

    @staticmethod
    def _rows(filename):
        with open(filename, "r") as f:
            return [row for row in csv.reader(f, delimiter=",")]

    def _finish_rewards(self):
        self.Rmax = float(self.R.max())
        self.Rmin = float(self.R.min())
        self.epsilon = (self.Rmax - self.Rmin) / 1000.0

    def load_raw_mdp(self, filename, scalarize=lambda x: x[0]):
        """Header n, ns, m, k, s0, ng, horizon, gamma; a row of ng goals; n*m rows of S
        (action-major: row a*n + s); n*m rows of T; then k blocks of m rows of R (row i*m + a
        holds R_i[., a] over the n states). scalarize maps the k reward arrays to one
        (default: the first). python/nova/file_loader.py:74-131."""
        d = self._rows(filename)
        try:
            h = d[0]
            n, ns, m, k = int(h[0]), int(h[1]), int(h[2]), int(h[3])
            self.n, self.ns, self.m = n, ns, m
            self.s0, self.ng = int(h[4]), int(h[5])
            self.horizon, self.gamma = int(h[6]), float(h[7])
            self.goals = np.array([int(x) for x in d[1][:self.ng]], dtype=np.int32)
            base = 2
            self.S = np.array([[[int(d[base + a * n + s][i]) for i in range(ns)]
                                for a in range(m)] for s in range(n)], dtype=np.int32)
            base += n * m
            self.T = np.array([[[float(d[base + a * n + s][i]) for i in range(ns)]
                                for a in range(m)] for s in range(n)], dtype=np.float32)
            base += n * m
            Rk = np.array([[[float(d[base + i * m + a][s]) for a in range(m)] for s in range(n)]
                           for i in range(k)], dtype=np.float32)
            self.R = np.asarray(scalarize(Rk), dtype=np.float32)
            self.k = k
        except (IndexError, ValueError) as e:
            raise ValueError("Failed to load file '%s': %s" % (filename, e))
        self._finish_rewards()

    def load_raw_pomdp(self, filename, scalarize=lambda x: x[0]):
        """Header n, ns, m, z, r, rz, k, s0, horizon, gamma; n*m rows of S; n*m rows of T; m*z
        rows of O (row a*z + o over the n successor states); k blocks of m rows of R; r rows
        of Z; r rows of B. python/nova/file_loader.py:133-213."""
        d = self._rows(filename)
        try:
            h = d[0]
            n, ns, m, z, r, rz, k = (int(x) for x in h[:7])
            self.n, self.ns, self.m, self.z, self.r, self.rz = n, ns, m, z, r, rz
            self.s0, self.horizon, self.gamma = int(h[7]), int(h[8]), float(h[9])
            base = 1
            self.S = np.array([[[int(d[base + a * n + s][i]) for i in range(ns)]
                                for a in range(m)] for s in range(n)], dtype=np.int32)
            base += n * m
            self.T = np.array([[[float(d[base + a * n + s][i]) for i in range(ns)]
                                for a in range(m)] for s in range(n)], dtype=np.float32)
            base += n * m
            self.O = np.array([[[float(d[base + a * z + o][sp]) for o in range(z)]
                                for sp in range(n)] for a in range(m)], dtype=np.float32)
            base += m * z
            Rk = np.array([[[float(d[base + i * m + a][s]) for a in range(m)] for s in range(n)]
                           for i in range(k)], dtype=np.float32)
            self.R = np.asarray(scalarize(Rk), dtype=np.float32)
            self.k = k
            base += k * m
            self.Z = np.array([[int(d[base + i][j]) for j in range(rz)] for i in range(r)],
                              dtype=np.int32)
            base += r
            self.B = np.array([[float(d[base + i][j]) for j in range(rz)] for i in range(r)],
                              dtype=np.float32)
        except (IndexError, ValueError) as e:
            raise ValueError("Failed to load file '%s': %s" % (filename, e))
        self._finish_rewards()

    # ------------------------------------------------------------------ Cassandra format

    @staticmethod
    def _statements(filename):
        """Splits the file into statements: a statement starts at a line holding ':' and runs
        until the next such line. '#' starts a comment; blank lines are dropped. Each statement
        is (head, body): the head line split at ':' and the following lines split into tokens."""
        out = []
        with open(filename, "r") as f:
            for raw in f:
                line = raw.split("#", 1)[0].strip()
                if not line:
                    continue
                if ":" in line:
                    out.append([line, []])
                elif out:
                    out[-1][1].append(line.split())
                else:
                    raise ValueError("data before the first statement: '%s'" % line)
        return out

    def load_cassandra(self, filename):
        """Reads a Cassandra-format POMDP (or MDP) file; python/nova/file_loader.py:215-587."""
        try:
            self._load_cassandra(self._statements(filename))
        except (IndexError, ValueError, KeyError) as e:
            raise ValueError("Failed to load file '%s': %s" % (filename, e))

    def _load_cassandra(self, stmts):
        header = {}
        # per kind, entries keyed by their field tuple: a repeated key keeps the position of its
        # first appearance and the value of its last (python/nova/file_loader.py:277-340 stores
        # them in a dict)
        entries = {"T": {}, "O": {}, "R": {}}
        for head, body in stmts:
            parts = [p.strip() for p in head.split(":")]
            name = parts[0]
            args = [p for p in parts[1:] if p]
            if name in entries:
                # fields: (action, state, next state / observation, ...); the last field may
                # carry the value(s) after its name
                last = args[-1].split() if args else []
                if not body and len(args) > 1:
                    key = tuple(args[:-1]) + (last[0],)
                    values = last[1:]
                else:
                    key = tuple(args)
                    values = body[0] if len(body) == 1 else body
                entries[name][key] = values
            else:
                header[name] = args[0].split() if args else (body[0] if len(body) == 1 else body)

        self.gamma = float(header["discount"][0]) if "discount" in header else 0.9

        def names(field):
            v = header.get(field)
            if v is None:
                return None
            if len(v) == 1 and v[0].isdigit():
                return [str(i) for i in range(int(v[0]))]
            return list(v)

        states, actions, observations = names("states"), names("actions"), names("observations")
        self.n = len(states) if states else 0
        self.m = len(actions) if actions else 0
        self.z = len(observations) if observations else 0
        n, m, z = self.n, self.m, self.z

        def pick(token, labels):
            return list(range(len(labels))) if token == "*" else [labels.index(token)]

        self._start(header, states)

        # T: successors appended in the order entries are met; only positive probabilities
        succ = [[[] for _ in range(m)] for _ in range(n)]
        prob = [[[] for _ in range(m)] for _ in range(n)]

        def add(s, a, sp, p):
            if p > 0.0:
                succ[s][a].append(sp)
                prob[s][a].append(p)

        def matrix(values):
            # one row given as a flat token list (n = 1) or rows of tokens
            return [values] if values and isinstance(values[0], str) else values

        for key, values in entries["T"].items():
            acts = pick(key[0], actions)
            if len(key) == 1:
                if values and values[0] == "uniform":
                    mat = [[1.0 / n] * n for _ in range(n)]
                elif values and values[0] == "identity":
                    mat = [[1.0 if s == sp else 0.0 for sp in range(n)] for s in range(n)]
                else:
                    mat = [[float(x) for x in row] for row in matrix(values)]
                for a in acts:
                    for s in range(n):
                        for sp in range(n):
                            add(s, a, sp, mat[s][sp])
            elif len(key) == 2:
                row = [1.0 / n] * n if values[0] == "uniform" else [float(x) for x in values]
                for a in acts:
                    for s in pick(key[1], states):
                        for sp in range(n):
                            add(s, a, sp, row[sp])
            else:
                p = float(values[0])
                for a in acts:
                    for s in pick(key[1], states):
                        for sp in pick(key[2], states):
                            add(s, a, sp, p)
        self.ns = max([1] + [len(succ[s][a]) for s in range(n) for a in range(m)])
        self.S = np.full((n, m, self.ns), -1, dtype=np.int32)
        self.T = np.zeros((n, m, self.ns), dtype=np.float32)
        for s in range(n):
            for a in range(m):
                c = len(succ[s][a])
                self.S[s, a, :c] = succ[s][a]
                self.T[s, a, :c] = prob[s][a]

        # O[a, sp, o]: assignments, later ones overriding
        self.O = np.zeros((m, n, z), dtype=np.float32)
        for key, values in entries["O"].items():
            acts = pick(key[0], actions)
            if len(key) == 1:
                if values and values[0] == "uniform":
                    mat = [[1.0 / z] * z for _ in range(n)]
                else:
                    mat = [[float(x) for x in row] for row in matrix(values)]
                for a in acts:
                    self.O[a] = np.asarray(mat, dtype=np.float32)
            elif len(key) == 2:
                row = [1.0 / z] * z if values[0] == "uniform" else [float(x) for x in values]
                for a in acts:
                    for sp in pick(key[1], states):
                        self.O[a, sp] = row
            else:
                p = float(values[0])
                for a in acts:
                    for sp in pick(key[1], states):
                        for o in pick(key[2], observations):
                            self.O[a, sp, o] = p

        # R[s, a] = mean of the entry's value(s); later entries override earlier ones
        self.k = 1
        self.R = np.zeros((n, m), dtype=np.float32)
        for key, values in entries["R"].items():
            flat = [float(x) for x in (values if not values or isinstance(values[0], str)
                                       else [y for row in values for y in row])]
            v = float(np.mean(np.asarray(flat, dtype=np.float64)))
            for a in pick(key[0], actions):
                for s in pick(key[1], states) if len(key) > 1 else range(n):
                    self.R[s, a] = v
        self._finish_rewards()
        # the horizon that brings the values within epsilon of optimal
        # (python/nova/file_loader.py:582-584); gamma < 1 is required
        if self.Rmax > self.Rmin and 0.0 < self.gamma < 1.0:
            self.horizon = int(math.log(self.epsilon / (self.Rmax - self.Rmin)) /
                               math.log(self.gamma)) + 1

    def _start(self, header, states):
        """The initial belief (one row of Z/B): 'start' as 'uniform', one state name, or n
        probabilities (kept as a full-length row, zeros included); or 'start include' /
        'start exclude' lists, uniform over the states they leave.
        python/nova/file_loader.py:386-432. ('start: uniform' is read as the uniform belief;
        the reference compares the token list with the string there and fails on it.)"""
        n = self.n
        # precedence as in the reference, whose later blocks override earlier ones:
        # start exclude, then start include, then start
        if "start exclude" in header:
            ex = set(header["start exclude"])
            idx = [s for s in range(n) if states[s] not in ex]
            self.Z = np.array([idx], dtype=np.int32)
            self.B = np.full((1, len(idx)), 1.0 / len(idx), dtype=np.float32)
        elif "start include" in header:
            idx = [states.index(x) for x in header["start include"]]
            self.Z = np.array([idx], dtype=np.int32)
            self.B = np.full((1, len(idx)), 1.0 / len(idx), dtype=np.float32)
        elif "start" in header:
            v = header["start"]
            if len(v) == 1 and v[0] == "uniform":
                self.Z = np.arange(n, dtype=np.int32)[None, :]
                self.B = np.full((1, n), 1.0 / n, dtype=np.float32)
            elif len(v) == 1:
                self.Z = np.array([[states.index(v[0])]], dtype=np.int32)
                self.B = np.ones((1, 1), dtype=np.float32)
            else:
                self.Z = np.arange(n, dtype=np.int32)[None, :]
                self.B = np.array([[float(x) for x in v]], dtype=np.float32)
        else:
            self.Z = self.B = None
        if self.Z is not None:
            self.r, self.rz = self.Z.shape
