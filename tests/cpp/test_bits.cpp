#include "bits.hpp"
#include <cstdio>
#include <random>
using namespace gfz;
int main(){
  std::mt19937_64 rng(1);
  long bad=0;
  for(int it=0;it<2000000;++it){
    uint64_t c = rng();
    int mode = it%4; if(mode==1) c |= rng(); if(mode==2) c &= rng(); if(mode==3) c = ~0ull >> (rng()%64);
    for(uint32_t p=0;p<2;++p){
      uint64_t ref=0; uint32_t prev=p;
      for(int i=0;i<64;++i){ uint32_t pk = ((c>>i)&1) && !prev; if(pk) ref |= 1ull<<i; prev=pk; }
      if(ref!=word_picks(c,p)) { if(bad<5) printf("bad c=%016llx p=%u\n",(unsigned long long)c,p); bad++; }
    }
  }
  // composition: random sequences of words
  for(int it=0;it<200000;++it){
    int n = 1 + rng()%8; uint64_t w[8]; for(int i=0;i<n;++i) w[i]=rng()&rng();
    for(uint32_t p=0;p<2;++p){
      uint32_t f = fn_identity(); for(int i=0;i<n;++i) f = fn_then(f, word_fn(w[i]));
      uint32_t prev=p, cnt=0; for(int i=0;i<n;++i) for(int b=0;b<64;++b){ uint32_t pk=((w[i]>>b)&1)&&!prev; cnt+=pk; prev=pk; }
      if(fn_out(f,p)!=prev || fn_cnt(f,p)!=cnt){ if(bad<10) printf("bad comp\n"); bad++; }
    }
  }
  printf("bad=%ld\n",bad); return bad!=0;
}
