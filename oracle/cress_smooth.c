/* 4th-order numerical smoothing (smtopt = 2: horizontal, 3: horizontal and
 * vertical) -- TEST INFRASTRUCTURE ONLY.  Included by cress_oracle.c.
 *
 * s_smoo4s (smoo4s.f90:171-300), s_smoo4qv (smoo4qv.f90), s_smoo4p
 * (smoo4p.f90) and s_smoo4uvw (smoo4uvw.f90:200-570) share one structure:
 * a prepared field a (rbr*s, rbr*(qv-qvbr), p', rst8*(u-ubr)/jcb8, ...),
 * its second differences t1 (x), t2 (y), t3 (z), the 2nd-order part
 * frc += smhcoe*(t1+t2) + smvcoe*t3 over the output box, and the 4th-order
 * part -- differences of the t's -- inside the band where the 4th-order
 * advection also applies (iwest..jnorth).  ex, ey, ez widen the ranges by
 * one for the staggered u (x), v (y) and w (z).  The reference keeps 2*a in
 * its own array; doubling is exact, so it is formed where used. */

static void or_smoo4_core(const or_cfg *c, int ex, int ey, int ez, const R *a, R *frc, R *t1, R *t2,
                          R *t3)
{
  DECL_DIMS(c);
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  const R smh = (R)c->smhcoe, smv = (R)c->smvcoe;
  const int khi = nk - 2 + ez;
  OMP_FOR
  for (int k = 2; k <= khi; k++) {
    for (int j = 2; j <= nj - 2 + ey; j++)
      for (int i = 1 + iw; i <= ni - 1 - ie + ex; i++)
        A3(t1, i, j, k) = (A3(a, i + 1, j, k) + A3(a, i - 1, j, k)) - K(2.0) * A3(a, i, j, k);
    for (int j = 1 + js; j <= nj - 1 - jn + ey; j++)
      for (int i = 2; i <= ni - 2 + ex; i++)
        A3(t2, i, j, k) = (A3(a, i, j + 1, k) + A3(a, i, j - 1, k)) - K(2.0) * A3(a, i, j, k);
    for (int j = 2; j <= nj - 2 + ey; j++)
      for (int i = 2; i <= ni - 2 + ex; i++)
        A3(t3, i, j, k) = (A3(a, i, j, k + 1) + A3(a, i, j, k - 1)) - K(2.0) * A3(a, i, j, k);
  }
  OMP_FOR
  for (int k = 2; k <= khi; k++)
    for (int j = 2; j <= nj - 2 + ey; j++)
      for (int i = 2; i <= ni - 2 + ex; i++)
        A3(frc, i, j, k) = A3(frc, i, j, k) + smh * (A3(t1, i, j, k) + A3(t2, i, j, k)) + smv * A3(t3, i, j, k);
  const int jlo = 2 + js, jhi = nj - 2 - jn + ey, ilo = 2 + iw, ihi = ni - 2 - ie + ex;
  if (c->smtopt % 10 == 2) {
    OMP_FOR
    for (int k = 2; k <= khi; k++)
      for (int j = jlo; j <= jhi; j++)
        for (int i = ilo; i <= ihi; i++)
          A3(frc, i, j, k) = A3(frc, i, j, k) +
                             smh * ((A3(t1, i, j, k) - (A3(t1, i + 1, j, k) + A3(t1, i - 1, j, k))) +
                                    (A3(t2, i, j, k) - (A3(t2, i, j + 1, k) + A3(t2, i, j - 1, k))));
    return;
  }
  for (int j = jlo; j <= jhi; j++)
    for (int i = ilo; i <= ihi; i++) {
      A3(t3, i, j, 1) = A3(t3, i, j, 2);
      A3(t3, i, j, khi + 1) = A3(t3, i, j, khi);
    }
  OMP_FOR
  for (int k = 2; k <= khi; k++)
    for (int j = jlo; j <= jhi; j++)
      for (int i = ilo; i <= ihi; i++)
        A3(frc, i, j, k) = A3(frc, i, j, k) +
                           (smv * (A3(t3, i, j, k) - (A3(t3, i, j, k + 1) + A3(t3, i, j, k - 1))) +
                            smh * ((A3(t1, i, j, k) - (A3(t1, i + 1, j, k) + A3(t1, i - 1, j, k))) +
                                   (A3(t2, i, j, k) - (A3(t2, i, j + 1, k) + A3(t2, i, j - 1, k)))));
}

/* The prepared field of a scalar over k=1..nk-1, j=jsouth..nj-jnorth,
 * i=iwest..ni-ieast: kind 0 rbr*s (smoo4s.f90:179-186), 1 rbr*(s-ref)
 * (smoo4qv.f90:176-183), 2 s itself (smoo4p.f90:166-172). */
void or_smoo4s(const or_cfg *c, int kind, const R *rbr, const R *ref, const R *s, R *frc, R *a,
               R *t1, R *t2, R *t3)
{
  DECL_DIMS(c);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = c->jsouth; j <= nj - c->jnorth; j++)
      for (int i = c->iwest; i <= ni - c->ieast; i++)
        A3(a, i, j, k) = kind == 0 ? A3(rbr, i, j, k) * A3(s, i, j, k)
                         : kind == 1 ? A3(rbr, i, j, k) * (A3(s, i, j, k) - A3(ref, i, j, k))
                                     : A3(s, i, j, k);
  or_smoo4_core(c, 0, 0, 0, a, frc, t1, t2, t3);
}

/* s_smoo4uvw: a = rst8u*(u-ubr)/jcb8u over i=iwest..ni+1-ieast
 * (smoo4uvw.f90:205-212), then v (j to nj+1-jnorth) and w
 * (rst8w*w/jcb8w, k=1..nk). */
void or_smoo4uvw(const or_cfg *c, const R *jcb8u, const R *jcb8v, const R *jcb8w, const R *ubr,
                 const R *vbr, const R *rst8u, const R *rst8v, const R *rst8w, const R *u, const R *v,
                 const R *w, R *ufrc, R *vfrc, R *wfrc, R *a, R *t1, R *t2, R *t3)
{
  DECL_DIMS(c);
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = js; j <= nj - jn; j++)
      for (int i = iw; i <= ni + 1 - ie; i++)
        A3(a, i, j, k) = A3(rst8u, i, j, k) * (A3(u, i, j, k) - A3(ubr, i, j, k)) / A3(jcb8u, i, j, k);
  or_smoo4_core(c, 1, 0, 0, a, ufrc, t1, t2, t3);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = js; j <= nj + 1 - jn; j++)
      for (int i = iw; i <= ni - ie; i++)
        A3(a, i, j, k) = A3(rst8v, i, j, k) * (A3(v, i, j, k) - A3(vbr, i, j, k)) / A3(jcb8v, i, j, k);
  or_smoo4_core(c, 0, 1, 0, a, vfrc, t1, t2, t3);
  OMP_FOR
  for (int k = 1; k <= nk; k++)
    for (int j = js; j <= nj - jn; j++)
      for (int i = iw; i <= ni - ie; i++)
        A3(a, i, j, k) = A3(rst8w, i, j, k) * A3(w, i, j, k) / A3(jcb8w, i, j, k);
  or_smoo4_core(c, 0, 0, 1, a, wfrc, t1, t2, t3);
}
