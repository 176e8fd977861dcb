/* species.h -- the species record passed across the collision-operator ABI.
 *
 * Layout identical to the reference's src/species.h:11-26: initialize_weights
 * takes it by value and initialize_conservation by pointer, so a driver built
 * against the reference headers passes it unchanged.
 */
#ifndef _SPECIES_H
#define _SPECIES_H

#include <stdlib.h>

typedef struct {
  size_t id;         /* identifier (macroscopic) */
  size_t num_levels; /* number of internal energy levels */
  size_t *lev_id;    /* global ids of the internal energy levels */
  double Rgas;       /* gas constant [J/kg] */
  double mass;       /* mass [kg] */
  double mm;         /* molecular mass [kg/mol] */
  double d_ref;      /* reference diameter [m] */
  double T_ref;      /* reference temperature [K] */
  double mu_ref;     /* reference dynamic viscosity [N s/m^2] */
  double omega;      /* viscosity-law temperature exponent */
  double E0;         /* formation energy [J] */
  double *Ei;        /* internal energy levels [J] */
  double *gi;        /* internal energy level degeneracies */
  char name[80];     /* name ("default" selects the old weight-file naming) */
} species;

#endif
