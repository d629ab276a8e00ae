// ref_capi.cpp -- exposes the reference's own `class Model` through the same C
// shape as the oracle (ora_*), so tests can compare the two bit for bit.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile from the reference sources
// where they lie under /root/reference (the file to include is passed as
// NICR_REF_SOURCE, the dimension as NICR_REF_DIM), with the reference's own
// single-process MPI shim (-DNICR_SERIAL_VERIFY, mpi2d/tests/mpi_serial.h).
// The same pattern is used by the reference's own drivers
// (mpi3d-slab/tests/regression_driver.cpp:1-8). The build output goes to
// oracle/_ref/ only; nothing here is part of the product.
#define main nicr_reference_main_unused
#include NICR_REF_SOURCE
#undef main

#include "../include/nicr.h"

namespace {
thread_local std::string g_ref_error;
struct Handle { std::unique_ptr<Model> m; std::string message; };

Config to_config(const nicr_config& n){
    Config c;
    c.nx=n.nx;c.ny=n.ny;
#if NICR_REF_DIM==3
    c.nz=n.nz;c.dz=n.dz;c.frozen_eta_cache=n.frozen_eta_cache!=0;
#endif
    c.steps=n.steps;c.output_every=n.output_every;c.left_edge_sites=n.left_edge_sites;
    c.random_seed=n.random_seed;c.dx=n.dx;c.dy=n.dy;c.dt=n.dt;
    c.grain_diameter_grid=n.grain_diameter_grid;c.gb_fwhm_grid=n.gb_fwhm_grid;
    c.gb_factor_o=n.gb_factor_o;c.gb_factor_cr=n.gb_factor_cr;
    c.radius_grid=n.radius_grid;c.exclusion_gap_grid=n.exclusion_gap_grid;c.nucleation_period=n.nucleation_period;
    c.mu_res=n.mu_res;c.surface_c_ref=n.surface_c_ref;
    c.bc=n.bc==NICR_BC_MU?"mu":(n.bc==NICR_BC_CLOSED?"closed":"concentration");
    c.ck=n.ck==NICR_CK_OFF?"off":"legacy";
    c.bounds=n.bounds==NICR_BOUNDS_LEGACY?"legacy":"stop";
    c.nucleation=n.nucleation==NICR_NUC_CENTER?"center":(n.nucleation==NICR_NUC_OFF?"off":"ksp");
    c.grains=n.grains!=0;c.oxygen_diffusion=n.oxygen_diffusion!=0;c.grain_evolution=n.grain_evolution!=0;
    c.eta_W=n.eta_W;c.eta_kappa=n.eta_kappa;c.eta_L=n.eta_L;c.eta_oxide_penalty=n.eta_oxide_penalty;
    c.max_eta_fields=n.max_eta_fields;
    return c;
}
#if NICR_REF_DIM==3
#define AT(f,i,j,k) (f)(i,j,k)
#define FOR_CELLS(m) for(int i=1;i<=(m).c.nx;++i)for(int j=1;j<=(m).c.ny;++j)for(int k=1;k<=(m).c.nz;++k)
#else
#define AT(f,i,j,k) (f)(i,j)
#define FOR_CELLS(m) for(int i=1;i<=(m).c.nx;++i)for(int j=1;j<=(m).c.ny;++j)for(int k=0;k<=0;++k)
#endif
Field* field_of(Model& m,int id){
    switch(id){
    case NICR_FIELD_O:return &m.o;case NICR_FIELD_CR:return &m.cr;case NICR_FIELD_PHI:return &m.p;
    case NICR_FIELD_GB:return &m.gb;case NICR_FIELD_MU_O:return &m.muO;case NICR_FIELD_MU_CR:return &m.muCr;
    case NICR_FIELD_M_O:return &m.mo;case NICR_FIELD_M_CR:return &m.mc;
    case NICR_FIELD_ETA_S2:return m.eta_s2.a.empty()?nullptr:&m.eta_s2;
#if NICR_REF_DIM==2
    case NICR_FIELD_ETA_PAIRS:return &m.eta_pairs;
#endif
    case 32:return m.grain.a.empty()?nullptr:&m.grain;
    default:
        if(id>=NICR_FIELD_ETA0 && id<NICR_FIELD_ETA0+int(m.eta.size()))return &m.eta[std::size_t(id-NICR_FIELD_ETA0)];
        return nullptr;
    }
}
void split(long double v,double out[2]){out[0]=double(v);out[1]=double(v-(long double)out[0]);}
long double join(const double in[2]){return (long double)in[0]+(long double)in[1];}
}  // namespace

extern "C" {
const char* ref_last_error(void){return g_ref_error.c_str();}
int ref_create(const nicr_config* cfg,void** out){
    try{auto* h=new Handle;h->m=std::make_unique<Model>(to_config(*cfg));*out=h;}
    catch(const std::exception& e){g_ref_error=e.what();return NICR_E_INVALID;}
    return NICR_OK;
}
int ref_destroy(void* h){delete static_cast<Handle*>(h);return NICR_OK;}
int ref_n_eta(void* h){return int(static_cast<Handle*>(h)->m->eta.size());}
int ref_n_seeds(void* h){return int(static_cast<Handle*>(h)->m->seeds.size());}
int ref_get_seeds(void* h,double* xyz){
    auto& m=*static_cast<Handle*>(h)->m;std::size_t q=0;
    for(const auto& s:m.seeds){xyz[q++]=s.x;xyz[q++]=s.y;
#if NICR_REF_DIM==3
        xyz[q++]=s.z;
#else
        xyz[q++]=0.;
#endif
    }
    return NICR_OK;
}
int ref_center_seed_cells(void* h){return static_cast<Handle*>(h)->m->center_seed_cells;}
int ref_get_initial(void* h,double out[4]){
    auto& m=*static_cast<Handle*>(h)->m;split(m.initialO,out);split(m.initialCr,out+2);return NICR_OK;
}
int ref_download_field(void* h,int id,double* dst){
    auto& m=*static_cast<Handle*>(h)->m;Field* f=field_of(m,id);
    if(!f){g_ref_error="unknown field";return NICR_E_INVALID;}
    std::size_t q=0;FOR_CELLS(m)dst[q++]=AT(*f,i,j,k);return NICR_OK;
}
int ref_upload_field(void* h,int id,const double* src){
    auto& m=*static_cast<Handle*>(h)->m;Field* f=field_of(m,id);
    if(!f){g_ref_error="unknown field";return NICR_E_INVALID;}
    std::size_t q=0;FOR_CELLS(m)AT(*f,i,j,k)=src[q++];
#if NICR_REF_DIM==3
    m.invalidate_eta_cache();
#endif
    return NICR_OK;
}
int ref_get_budget(void* h,nicr_budget* b){
    auto& m=*static_cast<Handle*>(h)->m;
    split(m.b.nucO,b->nucO);split(m.b.nucCr,b->nucCr);split(m.b.ckO,b->ckO);split(m.b.ckCr,b->ckCr);
    split(m.b.clipO,b->clipO);split(m.b.clipCr,b->clipCr);split(m.b.bcO,b->bcO);split(m.b.bcCr,b->bcCr);
    b->nuclei=m.b.nuclei;b->clipped=m.b.clipped;return NICR_OK;
}
int ref_set_budget(void* h,const nicr_budget* b){
    auto& m=*static_cast<Handle*>(h)->m;
    m.b.nucO=join(b->nucO);m.b.nucCr=join(b->nucCr);m.b.ckO=join(b->ckO);m.b.ckCr=join(b->ckCr);
    m.b.clipO=join(b->clipO);m.b.clipCr=join(b->clipCr);m.b.bcO=join(b->bcO);m.b.bcCr=join(b->bcCr);
    m.b.nuclei=b->nuclei;m.b.clipped=b->clipped;return NICR_OK;
}
int ref_refresh(void* h){static_cast<Handle*>(h)->m->refresh();return NICR_OK;}
// Model::step with the run loop's budget rollback (2D:1224-1225, SL:1373-1374).
// Returns NICR_STEP_*; the exception text is kept for ref_message().
int ref_step(void* h,int nsteps,int* done){
    auto* hd=static_cast<Handle*>(h);auto& m=*hd->m;int n=0;
    for(;n<nsteps;++n){
        Budget before=m.b;
        try{m.step();}
        catch(const std::exception& e){
            m.b=before;hd->message=e.what();if(done)*done=n;
            const std::string s=hd->message;
            if(s.rfind("Nonfinite",0)==0)return NICR_STEP_NONFINITE;
            if(s.rfind("Out-of-bounds",0)==0)return NICR_STEP_OUT_OF_BOUNDS;
            if(s.rfind("Eta trial",0)==0)return NICR_STEP_ETA_RANGE;
            return -100;
        }
    }
    if(done)*done=n;
    return NICR_STEP_OK;
}
const char* ref_message(void* h){return static_cast<Handle*>(h)->message.c_str();}
int ref_nucleate(void* h){static_cast<Handle*>(h)->m->nucleate();return NICR_OK;}
int ref_stamp(void* h,int x,int y,int z){
#if NICR_REF_DIM==3
    return static_cast<Handle*>(h)->m->stamp(x,y,z);
#else
    (void)z;return static_cast<Handle*>(h)->m->stamp(x,y);
#endif
}
int ref_valid_nucleus(void* h,int x,int y,int z){
#if NICR_REF_DIM==3
    return static_cast<Handle*>(h)->m->valid_nucleus(x,y,z)?1:0;
#else
    (void)z;return static_cast<Handle*>(h)->m->valid_nucleus(x,y)?1:0;
#endif
}
int ref_sum(void* h,int id,double out[2]){
    auto& m=*static_cast<Handle*>(h)->m;Field* f=field_of(m,id);
    if(!f){g_ref_error="unknown field";return NICR_E_INVALID;}
    split(Model::sum(*f),out);return NICR_OK;
}
// The reference's own writers, for byte comparisons with the drop-in writers.
// what: 0 metadata into directory `path`, 1 diagnostics row, 2 grain rows
// ("diag|history"), 3 legacy VTK, 4 VTI (3D only). Rows are appended.
int ref_write(void* h,int what,const char* path,int step,const char* fmt){
    auto& m=*static_cast<Handle*>(h)->m;
    try{
        if(what==0){m.c.out=path;m.metadata();}
        else if(what==1){std::ofstream f(path,std::ios::app);m.log(f,step);}
        else if(what==2){
            const std::string s(path);const auto bar=s.find('|');
            std::ofstream f(s.substr(0,bar),std::ios::app),g(s.substr(bar+1),std::ios::app);
            m.grain_log(f,g,step);
        }else if(what==3){m.vtk(path);}
#if NICR_REF_DIM==3
        else if(what==4){if(fmt&&*fmt)m.c.output_precision=fmt;m.vti(path);}
#endif
        else{g_ref_error="unknown writer";return NICR_E_INVALID;}
    }catch(const std::exception& e){g_ref_error=e.what();return NICR_E_INVALID;}
    (void)fmt;return NICR_OK;
}
// Timing helper for the CPU baseline: nsteps plain Model::step calls.
int ref_run(void* h,int nsteps){
    auto& m=*static_cast<Handle*>(h)->m;
    for(int n=0;n<nsteps;++n){try{m.step();}catch(const std::exception&){return n;}}
    return nsteps;
}
}  // extern "C"
