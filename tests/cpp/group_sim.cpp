// Host check of the formulation behind group_children (paulib_amd/csrc/
// plh_collate.hip): the children of a split bucket are groups of its 256
// sub-buckets, defined sequentially as a greedy walk (large sub-buckets
// alone, runs of small ones merged while their total fits the cap, empty
// ones skipped).  The kernel computes them in parallel instead: each
// sub-bucket d gets the end nxt(d) of a group that would start there (a
// binary search on the size prefix, capped by the next large sub-bucket; a
// large one ends at the next non-empty one), and the groups formed are the
// sub-buckets reached from the first non-empty one by following nxt, found
// with 8 rounds of pointer doubling.  This program restates both and
// compares them on 200,000 seeded size vectors of several shapes (sparse,
// dense, with and without large sub-buckets, all empty).  Exit status 0
// and "0/200000 mismatches" on success.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>
#include <tuple>
using G = std::tuple<uint64_t, uint64_t, uint32_t>;
std::vector<G> seq(const uint64_t* st, const uint64_t* ss, uint32_t cap) {
  std::vector<G> out; uint64_t g0=0,gn=0; uint32_t gc=0;
  for (int d=0; d<256; ++d) { uint64_t sz=ss[d]; if(!sz) continue;
    if (sz>cap){ if(gn){out.emplace_back(g0,gn,gc);gn=0;gc=0;} out.emplace_back(st[d],sz,1u); continue;}
    if (gn+sz>cap){out.emplace_back(g0,gn,gc);gn=0;gc=0;}
    if(!gn) g0=st[d]; gn+=sz; ++gc; }
  if (gn) out.emplace_back(g0,gn,gc);
  return out;
}
std::vector<G> par(const uint64_t* st, const uint64_t* ss, uint32_t cap) {
  uint64_t P[257]; uint32_t NZ[257]; uint32_t J[257], M[257], nxt[256];
  P[0]=0; NZ[0]=0;
  for (int d=0; d<256; ++d){P[d+1]=P[d]+ss[d]; NZ[d+1]=NZ[d]+(ss[d]!=0);}
  int c0=-1; for(int d=0; d<256; ++d) if(ss[d]){c0=d;break;}
  std::vector<G> out; if (c0<0) return out;
  // next big strictly after d
  int nbA[256], nzA[256]; int nb=256, nz=256; for(int d=255; d>=0; --d){ nbA[d]=nb; nzA[d]=nz; if(ss[d]>cap) nb=d; if(ss[d]) nz=d; }
  for (int d=0; d<256; ++d) {
    if (ss[d]>cap) { nxt[d]=nzA[d]; }
    else {
      uint64_t t=P[d]+cap; uint32_t lo=d+1, n=257-lo;
      while(n>0){ uint32_t half=n/2, m=lo+half; if(P[m]<=t){lo=m+1;n-=half+1;} else n=half; }
      uint32_t ub=lo; uint32_t e = ub-1; if (ub>256) e=256;
      nxt[d] = (uint32_t)nbA[d] < e ? (uint32_t)nbA[d] : e;
    }
    J[d]=nxt[d]; M[d]=0;
  }
  J[256]=256; M[256]=0; M[c0]=1;
  for (int r=0;r<8;++r){ uint32_t j[256], jj[256], m[256];
    for(int d=0;d<256;++d){j[d]=J[d];jj[d]=J[J[d]];m[d]=M[d];}
    for(int d=0;d<256;++d){ if(m[d]) M[j[d]]=1; J[d]=jj[d]; } }
  for (int d=0; d<256; ++d) if (M[d]) out.emplace_back(st[d], P[nxt[d]]-P[d], NZ[nxt[d]]-NZ[d]);
  return out;
}
int main(){ std::mt19937_64 rng(7); long bad=0, tot=0;
  for (int it=0; it<200000; ++it){ uint64_t ss[256], st[256]; uint32_t cap = 1 + rng()%3000;
    int mode = rng()%6; uint64_t s0 = rng()%1000000;
    for (int d=0; d<256; ++d){ uint64_t v;
      switch(mode){case 0: v = rng()%4==0 ? 0 : rng()%(cap/4+1); break;
        case 1: v = rng()%2 ? 0 : rng()%(2*cap+2); break;
        case 2: v = rng()%10==0 ? rng()%(cap+1) : 0; break;
        case 3: v = rng()%(cap+1); break;
        case 4: v = 0; break;
        default: v = (rng()%50==0) ? (uint64_t)cap + rng()%100000 : (rng()%3 ? 0 : rng()%(cap/16+1)); }
      if (mode==4 && rng()%200==0) v = rng()%(3*cap);
      ss[d]=v; st[d]=s0; s0+=v; }
    auto a=seq(st,ss,cap), b=par(st,ss,cap); ++tot; if (a!=b){ if(++bad<5) printf("mismatch it=%d mode=%d cap=%u sizes %zu %zu\n",it,mode,cap,a.size(),b.size()); } }
  printf("%ld/%ld mismatches\n", bad, tot); return bad!=0; }
