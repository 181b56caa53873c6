// material.cpp -- RNPBond material constants on the host.
//
// Mirrors material::pd::RNPBond::RNPBond (src/material/pd/rnpBond.cpp:23-97)
// with computeParameters (:99-189) / computeMaterialProperties (:191-212),
// the influence functions of src/material/pd/influenceFn.h:85-408 and the
// elastic conversions of inp::MatData (src/inp/decks/materialDeck.h:112-215).
// The expressions keep the reference's order so the constants are the same
// doubles the reference computes.

#include <cmath>
#include <cstring>

#include "nlm_internal.h"

namespace {

struct InfFn {
  int type;
  double a, b;
  // getMoment(i): ConstInfluenceFn :117-119, LinearInfluenceFn, Gaussian
  double moment(int i) const {
    if (type == 0) return a / double(i + 1);
    if (type == 1) return (a / double(i + 1)) + (b / double(i + 2));
    const double sq1 = std::sqrt(b), sq2 = std::sqrt(M_PI);
    switch (i) {
      case 0:
        return 0.5 * a * sq1 * sq2 * std::erf(1. / sq1);
      case 1:
        return 0.5 * a * b * (1. - std::exp(-1. / b));
      case 2:
        return 0.5 * a * b * sq1 *
               (0.5 * sq2 * std::erf(1. / sq1) -
                (1. / sq1) * std::exp(-1. / b));
      default:
        return 0.5 * a * b * b * (1. - (1. + 1. / b) * std::exp(-1. / b));
    }
  }
};

// Influence-function parameters with the reference defaults.
bool make_inf(const nlm_material_deck* d, InfFn* f) {
  f->type = d->inf_type;
  const int np = d->n_inf_params;
  const double* p = d->inf_params;
  const int dim = d->dim;
  switch (d->inf_type) {
    case 0:  // constant, default a0 = dim + 1 (influenceFn.h:95)
      f->a = np > 0 ? p[0] : double(dim + 1);
      f->b = 0.;
      return true;
    case 1:  // linear a0 + a1 r, default a0 (1 - r) normalised
      if (np == 0) {
        f->a = dim == 1 ? 6. : (dim == 2 ? 12. : 20.);
        f->b = -f->a;
      } else {
        f->a = p[0];
        f->b = np < 2 ? -f->a : p[1];
      }
      return true;
    case 2:  // gaussian alpha exp(-r^2/beta), default beta = 0.2
      if (np == 0) {
        const double beta = 0.2;
        f->b = beta;
        if (dim == 1)
          f->a = 2. / (beta * (1. - std::exp(-1. / beta)));
        else if (dim == 2)
          f->a = (4.0 / beta) * 1.0 /
                 (std::sqrt(M_PI * beta) * std::erf(1.0 / std::sqrt(beta)) -
                  2.0 * std::exp(-1.0 / beta));
        else
          f->a = (2.0 / beta) * 1.0 /
                 (beta - (beta + 1.) * std::exp(-1.0 / beta));
      } else {
        if (np < 2) return false;
        f->a = p[0];
        f->b = p[1];
      }
      return true;
    default:
      return false;
  }
}

// definitelyLessThan (src/util/compare.cpp:30-33)
bool less_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (b - a) > m * 1e-5;
}

}  // namespace

extern "C" int nlm_rnp_params_from_deck(const nlm_material_deck* d,
                                        nlm_rnp_params* o) {
  if (!d || !o) return NLM_ERR_INVALID;
  if (d->dim < 1 || d->dim > 3 || !(d->horizon > 0.)) return NLM_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  InfFn J;
  if (!make_inf(d, &J)) return NLM_ERR_INVALID;
  o->dim = d->dim;
  o->horizon = d->horizon;
  o->h = d->h;
  o->density = d->density;
  o->inf_type = J.type;
  o->inf_a = J.a;
  o->inf_b = J.b;
  const double M = J.moment(d->dim);
  o->inf_moment = M;
  o->irreversible = d->irreversible ? 1 : 0;
  o->factor_sc = d->check_sc_factor;
  switch (d->dim) {  // rnpBond.cpp:67-72
    case 1:
      o->inv_factor = std::pow(d->horizon, 2) * 2.;
      break;
    case 2:
      o->inv_factor = std::pow(d->horizon, 3) * M_PI;
      break;
    default:
      o->inv_factor = std::pow(d->horizon, 4) * 4. * M_PI / 3.;
      break;
  }
  double E = d->E, K = d->K, Gc = d->Gc, KIc = d->KIc;
  double nu, lambda, G;
  if (d->compute_from_classical) {
    if (less_than(E, 0.) && less_than(K, 0.)) return NLM_ERR_INVALID;
    if (less_than(Gc, 0.) && less_than(KIc, 0.)) return NLM_ERR_INVALID;
    nu = d->plane_strain ? 0.25 : 1. / 3.;  // rnpBond.cpp:138-153
    if (E > 0.) K = E / (3. * (1. - 2. * nu));               // toK
    if (K > 0. && E < 0.) E = K * (3. * (1. - 2. * nu));     // toE
    if (Gc > 0.) KIc = std::sqrt(Gc * E);                    // toKIc
    if (KIc > 0. && Gc < 0.) Gc = KIc * KIc / E;             // toGc
    lambda = E * nu / ((1. + nu) * (1. - 2. * nu));          // toLambdaE
    G = E / (2. * (1. + nu));                                // toGE
    if (d->dim == 2) {
      o->C = M_PI * Gc / (4. * M);
      o->beta = 4. * lambda / (o->C * M);
    } else if (d->dim == 3) {
      o->C = 2. * Gc / (3. * M);
      o->beta = 5. * lambda / (o->C * M);
    }
    o->rbar = std::sqrt(0.5 / o->beta);
  } else {
    if (d->n_bond_params < 2) return NLM_ERR_INVALID;
    o->C = d->bond_params[0];
    o->beta = d->bond_params[1];
    // rbar = sqrt(0.5 / beta) and the bond's exponent exp(-beta r S^2) need
    // beta > 0 (the reference computes NaN / inf force laws otherwise; the
    // FP32-force staged kernel scales displacements by sqrt(beta log2 e))
    if (!(o->beta > 0.) || !std::isfinite(o->beta)) return NLM_ERR_INVALID;
    o->rbar = std::sqrt(0.5 / o->beta);
    nu = 0.33;  // computeMaterialProperties, rnpBond.cpp:191-212
    if (d->dim == 2) {
      Gc = 4. * M * o->C / M_PI;
      lambda = o->C * M * o->beta / 4.;
    } else {
      Gc = 3. * M * o->C / 2.;
      lambda = o->C * M * o->beta / 5.;
    }
    G = lambda;
    E = lambda * (1. + nu) * (1. - 2. * nu) / nu;
    K = E / (3. * (1. - 2. * nu));
    KIc = std::sqrt(Gc * E);
  }
  o->E = E;
  o->K = K;
  o->nu = nu;
  o->lambda = lambda;
  o->G = G;
  o->Gc = Gc;
  o->KIc = KIc;
  o->contact = d->no_penetration ? 1 : 0;
  if (o->contact) {  // rnpBond.cpp:85-94
    o->contact_rc = 0.9 * d->h;
    o->contact_kn = K * 18. / (M_PI * std::pow(d->horizon, 5));
  }
  return NLM_OK;
}

// ElasticState::ElasticState / computeParameters (ElasticState.cpp:17-158):
// the checks where the reference exits become NLM_ERR_INVALID.
extern "C" int nlm_es_params_from_deck(const nlm_es_deck* d,
                                       nlm_es_params* o) {
  if (!d || !o) return NLM_ERR_INVALID;
  if (d->dim < 1 || d->dim > 3) return NLM_ERR_INVALID;
  if (!(d->horizon > 0.) || !(d->h > 0.) || !(d->density > 0.))
    return NLM_ERR_INVALID;
  // 1D needs E (:61-66); 2D rejects E together with K and needs K and G
  // (:70-91); K and G drive the 3D constants
  if (d->dim == 1 && !(d->E > 0.)) return NLM_ERR_INVALID;
  if (d->dim == 2 && d->E > 0. && d->K > 0.) return NLM_ERR_INVALID;
  if (d->dim >= 2 && (!(d->K > 0.) || !(d->G > 0.))) return NLM_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  o->dim = d->dim;
  o->horizon = d->horizon;
  o->h = d->h;
  o->density = d->density;
  o->E = d->E;
  o->K = d->K;
  o->G = d->G;
  o->plane_strain = d->plane_strain ? 1 : 0;
  o->critical_stretch = d->critical_stretch;
  // MatData::toNuClassical (materialDeck.h:120), mu = G (:94-96)
  o->nu = (3 * d->K - 2 * d->G) / (2 * (3 * d->K + d->G));
  o->mu = d->G;
  if (!o->plane_strain && d->dim == 2)
    o->factor2D = (2. * o->nu - 1.) / (o->nu - 1.);
  else
    o->factor2D = 1;
  return NLM_OK;
}
