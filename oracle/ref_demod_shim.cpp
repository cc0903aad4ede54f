// extern "C" access to the reference's own soft demappers and pseudo-random generator,
// built from their sources (oracle/Makefile.ref) -- TEST INFRASTRUCTURE ONLY: the checker of
// the oracle's demodulation and scrambling restatement.  Built without AVX/NEON flags, so
// the reference's scalar (generic) demapper path runs.
#include "demodulation_mapper_impl.h"
#include "pseudo_random_generator_impl.h"
#include "ocudu/adt/bit_buffer.h"

#include <vector>

using namespace ocudu;

extern "C" {

// demodulation_mapper::demodulate_soft over nsym symbols (re, im interleaved).
void ref_demodulate_soft(int8_t* llrs, const float* iq, const float* nv, unsigned nsym, int modulation)
{
  static demodulation_mapper_impl demapper;
  const unsigned qm = modulation == 0 ? 1u : (unsigned)modulation;
  span<log_likelihood_ratio> out(reinterpret_cast<log_likelihood_ratio*>(llrs), (size_t)nsym * qm);
  span<const cf_t> sym(reinterpret_cast<const cf_t*>(iq), nsym);
  span<const float> n(nv, nsym);
  demapper.demodulate_soft(out, sym, n, static_cast<modulation_scheme>(modulation));
}

// pseudo_random_generator: init(c_init), advance(offset), generate(n bits); one bit per byte.
void ref_scrambling_sequence(uint8_t* c, uint32_t c_init, unsigned offset, unsigned n)
{
  pseudo_random_generator_impl prg;
  prg.init(c_init);
  prg.advance(offset);
  dynamic_bit_buffer seq(n);
  prg.generate(seq);
  for (unsigned i = 0; i < n; ++i) {
    c[i] = (uint8_t)seq.extract(i, 1);
  }
}

} // extern "C"
