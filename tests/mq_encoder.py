"""MQ arithmetic encoder (ITU-T T.800 C.2: INITENC, CODEMPS/CODELPS, RENORME,
BYTEOUT, FLUSH), written for the tests so that decoders can be checked on
symbol streams they did not produce themselves."""
import json
import os

_T = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_tables.json")))
QE, NMPS, NLPS, SWITCH = _T["Qe"], _T["NMPS"], _T["NLPS"], _T["SWITCH"]

# context labels as in the oracle: 0..8 ZC, 9..13 SC, 14..16 MR, 17 RL, 18 UNI
INITIAL = {0: 4, 17: 3, 18: 46}


class MQEncoder:
    def __init__(self, nctx=19):
        self.A, self.C, self.CT = 0x8000, 0, 12
        self.buf = [0]  # buf[0] is the byte before the codeword (BPST - 1)
        self.I = [INITIAL.get(i, 0) for i in range(nctx)]
        self.MPS = [0] * nctx

    def _byteout(self):
        if self.buf[-1] == 0xFF:
            self.buf.append(self.C >> 20)
            self.C &= 0xFFFFF
            self.CT = 7
        elif self.C < 0x8000000:
            self.buf.append(self.C >> 19)
            self.C &= 0x7FFFF
            self.CT = 8
        else:
            self.buf[-1] += 1
            if self.buf[-1] == 0xFF:
                self.C &= 0x7FFFFFF
                self.buf.append(self.C >> 20)
                self.C &= 0xFFFFF
                self.CT = 7
            else:
                self.buf.append((self.C >> 19) & 0xFF)  # the carry went into the previous byte
                self.C &= 0x7FFFF
                self.CT = 8

    def _renorm(self):
        while True:
            self.A <<= 1
            self.C <<= 1
            self.CT -= 1
            if self.CT == 0:
                self._byteout()
            if self.A & 0x8000:
                break

    def encode(self, d, cx):
        i = self.I[cx]
        qe = QE[i]
        self.A -= qe
        if d == self.MPS[cx]:
            if (self.A & 0x8000) == 0:
                if self.A < qe:
                    self.A = qe
                else:
                    self.C += qe
                self.I[cx] = NMPS[i]
                self._renorm()
            else:
                self.C += qe
        else:
            if self.A < qe:
                self.C += qe
            else:
                self.A = qe
            if SWITCH[i]:
                self.MPS[cx] ^= 1
            self.I[cx] = NLPS[i]
            self._renorm()

    def flush(self):
        tempc = self.C + self.A
        self.C |= 0xFFFF
        if self.C >= tempc:
            self.C -= 0x8000
        self.C <<= self.CT
        self._byteout()
        self.C <<= self.CT
        self._byteout()
        out = self.buf[1:]
        if out and out[-1] == 0xFF:
            out = out[:-1]
        return bytes(out)
