#include <cstdio>
#include <random>

// Host build of the bitsliced GF(2^32) constant multiplication used by the QC
// encoder (cuot_amd/csrc/gf32.h): both the wave-uniform and the per-lane forms
// against scalar field multiplication, 32 rows per trial.
#include "../cuot_amd/csrc/gf32.h"
int main(){
  std::mt19937 r(3);
  for (int trial=0; trial<2000; trial++){
    uint32_t v[32], acc[32], acc2[32]; for(int j=0;j<32;j++) v[j]=r();
    uint32_t c = trial<32 ? (1u<<trial) : r();
    gf32_mulc_top<true>(v,c,acc); gf32_mulc_top<false>(v,c,acc2);
    for (int row=0; row<32; row++){
      uint32_t x=0; for(int j=0;j<32;j++) x|=((v[j]>>row)&1u)<<j;
      uint32_t y=gf32_mul_host(c,x), y1=0, y2=0;
      for(int j=0;j<32;j++){ y1|=((acc[j]>>row)&1u)<<j; y2|=((acc2[j]>>row)&1u)<<j; }
      if(y!=y1||y!=y2){printf("FAIL trial %d row %d c=%08x\n",trial,row,c);return 1;}
    }
  }
  printf("gf32 top ok\n");
}
