"""Random UL-SCH multiplexing configurations (ulsch_demultiplex::configuration plus CSI
Part 2 sizes) for the demultiplexer tests: every modulation and layer count, one to three
DM-RS positions (single or double symbol), both DM-RS types, HARQ-ACK of 0..2 bits over
reserved REs (the reservation sized as get_ulsch_information does, for two bits, also
when there is no HARQ-ACK: lib/ran/pusch/ulsch_info.cpp:225-252) or of more bits without
reservation, CSI Part 1 and Part 2 of 1, 2 or more bits, field sizes from one RE to most
of the UCI REs.  Configurations the oracle rejects are drawn again."""
import numpy as np

import oracle as O


def _dmrs_mask(rng, start, nsym):
    first = start + int(rng.choice([0, 2])) if nsym > 3 else start
    mask = 1 << first
    if rng.random() < 0.3 and first + 1 < start + nsym - 1:
        mask |= 1 << (first + 1)  # double-symbol DM-RS
    for _ in range(int(rng.integers(0, 3))):
        l = int(rng.integers(first + 2, start + nsym)) if first + 2 < start + nsym else first
        mask |= 1 << l
    return mask


def random_config(rng, max_prb=24, uci=True, qm_layers=None):
    while True:
        mod = int(rng.choice([0, 1, 2, 4, 6, 8]))
        qm = 1 if mod == 0 else mod
        layers = 1 if mod == 0 else int(rng.integers(1, 5))
        if qm_layers is not None:
            mod, layers = qm_layers
            qm = 1 if mod == 0 else mod
        nbpr = qm * layers
        prb = int(rng.integers(1, max_prb + 1))
        start = int(rng.integers(0, 4))
        nsym = int(rng.integers(4, 15 - start))
        mask = _dmrs_mask(rng, start, nsym)
        dtype = int(rng.integers(1, 3))
        cdm = int(rng.integers(1, 3 if dtype == 1 else 4))
        cfg = dict(modulation=mod, nof_layers=layers, nof_prb=prb, start_symbol_index=start, nof_symbols=nsym,
                   dmrs_symbol_mask=mask, dmrs_type=dtype, nof_cdm_groups_without_data=cdm)
        if uci:
            n_data = sum(1 for l in range(start, start + nsym) if not (mask >> l) & 1)
            uci_re = n_data * prb * 12

            def res(frac):
                return max(1, int(uci_re * frac * rng.random()))

            ack = int(rng.choice([0, 0, 1, 2, 3, 7, 20]))
            if ack <= 2:
                rvd = res(0.3) if rng.random() < 0.8 else 0
                ack_re = 0 if ack == 0 or rvd == 0 else (rvd if ack == 2 else int(rng.integers(1, rvd + 1)))
                if ack and not ack_re:
                    ack = 0
            else:
                rvd, ack_re = 0, res(0.3)
            c1 = int(rng.choice([0, 1, 2, 9, 40]))
            c1_re = res(0.3) if c1 else 0
            c2 = int(rng.choice([0, 0, 1, 2, 60])) if c1 else 0
            c2_re = res(0.3) if c2 else 0
            cfg.update(nof_harq_ack_rvd=rvd * nbpr, nof_harq_ack_bits=ack, nof_enc_harq_ack_bits=ack_re * nbpr,
                       nof_csi_part1_bits=c1, nof_enc_csi_part1_bits=c1_re * nbpr, nof_csi_part2_bits=c2,
                       nof_enc_csi_part2_bits=c2_re * nbpr)
        if O.demux_sizes(cfg) is not None:
            return cfg


def config_words(cfg, c_init):
    return [int(cfg.get(f, 0)) for f in O.DEMUX_FIELDS] + [int(c_init)]
